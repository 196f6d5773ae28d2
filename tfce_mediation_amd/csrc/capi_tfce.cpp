// C-ABI host layer: adjacency handles, per-stream scratch, TFCE entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <unordered_map>
#include <vector>

#include "tfcehip.h"
#include "capi_common.h"
#include "tfce_internal.h"

using namespace tfcehip;

struct tfce_adj {
  int32_t V = 0;
  int64_t nnz = 0;
  float H = 2.0f, E = 1.0f;
  int32_t flags = 0;
  int device = 0;
  int32_t* d_row_ptr = nullptr;
  int32_t* d_col = nullptr;
  int32_t* d_ell = nullptr;     // [V][ell_width], -1 padded; null for long lists
  uint8_t* d_nadj = nullptr;    // [V][ell_width] adjacency among a row's neighbours (width <= 8)
  uint32_t* d_adjw = nullptr;   // [V or distinct rows][ell_width] full row adjacency, rows wider than kNposMaxWidth
  uint16_t* d_adjid = nullptr;  // [V] row of d_adjw per element when d_adjw holds distinct rows only
  int32_t ell_width = 0;
  double* d_sizepow = nullptr;  // (double)s^E for s in [0, V] when E != 1
  SurfDev* d_surf = nullptr;    // one descriptor, offset 0
};

struct tfce_surfset {
  int32_t n = 0;
  int device = 0;
  int32_t v_max = 0;
  int32_t np_w = 0;  // widest neighbour-position table among the surfaces
  std::vector<SurfDev> host;  // in launch order (largest first)
  std::vector<int32_t> user_index;  // launch position -> caller's surface index
  bool all_ring = true;             // every surface ring-packed
  std::vector<uint8_t> lds_ok;      // per launch position: the LDS-resident sweep may take it
  SurfDev* d_surfs = nullptr;
};

namespace tfcehip {

thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

static std::mutex g_ws_mutex;
static std::map<std::pair<int, hipStream_t>, Workspace> g_ws;
static int g_wgs_per_cu = 4;
static bool g_lds_sweep = true;  // small surfaces in the LDS-resident sweep
static bool g_pair_sweep = true;  // both signs of a map in one workgroup (ring-packed MAX sweeps)
// tfce_run on one image: -1 the whole-GPU sweep for symmetric surfaces of at
// least kCoopMinV elements, 0 never, 1 for every symmetric surface
static int g_coop_mode = -1;
static int g_coop_grid = 0;
constexpr int32_t kCoopMinV = 32768;

int current_device() {
  int d = 0;
  (void)hipGetDevice(&d);
  return d;
}

int cu_count(int device) {
  static std::map<int, int> cache;
  static std::mutex m;
  std::lock_guard<std::mutex> g(m);
  auto it = cache.find(device);
  if (it != cache.end()) return it->second;
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0)
    n = 256;
  cache[device] = n;
  return n;
}

size_t device_memory(int device) {
  static std::map<int, size_t> cache;
  static std::mutex m;
  std::lock_guard<std::mutex> g(m);
  auto it = cache.find(device);
  if (it != cache.end()) return it->second;
  size_t bytes = 0;
  if (hipDeviceTotalMem(&bytes, device) != hipSuccess || bytes == 0) bytes = (size_t)64 << 30;
  cache[device] = bytes;
  return bytes;
}

bool stream_capturing(hipStream_t stream) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cs) != hipSuccess) return false;
  return cs == hipStreamCaptureStatusActive;
}

static std::mutex g_retained_mutex;
static std::vector<void*> g_retained;  // superseded buffers a captured graph may use

void scratch_release(void* ptr, bool captured, hipStream_t stream) {
  if (!ptr) return;
  if (captured) {
    std::lock_guard<std::mutex> g(g_retained_mutex);
    g_retained.push_back(ptr);
    return;
  }
  (void)hipStreamSynchronize(stream);
  (void)hipFree(ptr);
}

// Grow-only scratch for (device, stream).  Growing synchronises the stream so
// no kernel of that stream still uses the old buffer.  A call made while the
// stream captures a graph never allocates (the scratch must have been grown
// by an eager call first) and marks the buffer as held by a graph: when a
// later call grows it, the old buffer is kept allocated, since a replay of
// that graph, on any stream, still writes it.
int workspace_get(hipStream_t stream, size_t bytes, Workspace** out) {
  const int dev = current_device();
  std::lock_guard<std::mutex> g(g_ws_mutex);
  Workspace& w = g_ws[std::make_pair(dev, stream)];
  if (stream_capturing(stream)) {
    if (!w.queue || w.bytes < bytes)
      return set_error(TFCE_EINVAL, "TFCE scratch of this stream is too small to capture this call: run it "
                                    "once eagerly on the capture stream first");
    w.captured = true;
    *out = &w;
    return 0;
  }
  if (!w.queue) {
    hipError_t e = hipMalloc(&w.queue, 256);
    if (e != hipSuccess) return set_error(-(int)e, std::string("hipMalloc(queue): ") + hipGetErrorString(e));
  }
  if (w.bytes < bytes) {
    if (w.ptr) {
      scratch_release(w.ptr, w.captured, stream);
      w.ptr = nullptr;
      w.bytes = 0;
      w.captured = false;
    }
    size_t want = std::max(bytes, (size_t)(w.bytes * 5 / 4));
    hipError_t e = hipMalloc(&w.ptr, want);
    if (e != hipSuccess) {
      e = hipMalloc(&w.ptr, bytes);
      if (e != hipSuccess)
        return set_error(-(int)e, std::string("hipMalloc(scratch): ") + hipGetErrorString(e));
      want = bytes;
    }
    w.bytes = want;
  }
  *out = &w;
  return 0;
}

int tfce_grid(int64_t n_items) {
  const int64_t slots = (int64_t)cu_count(current_device()) * g_wgs_per_cu;
  return (int)std::max<int64_t>(1, std::min<int64_t>(n_items, slots));
}

}  // namespace tfcehip

#define HIPCALL(expr)                                                               \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return set_error(-(int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

extern "C" {

const char* tfce_last_error(void) { return g_last_error.c_str(); }
const char* tfce_version(void) { return "tfcehip 0.1 gfx950"; }

int tfce_set_lds_sweep(int32_t enable) {
  g_lds_sweep = enable != 0;
  return 0;
}

int tfce_set_pair_sweep(int32_t enable) {
  g_pair_sweep = enable != 0;
  return 0;
}

int tfce_set_single_image_mode(int32_t mode, int32_t max_workgroups) {
  if (mode < -1 || mode > 1) return set_error(TFCE_EINVAL, "mode must be -1 (auto), 0 or 1");
  if (max_workgroups < 0) return set_error(TFCE_EINVAL, "max_workgroups must be >= 0");
  g_coop_mode = mode;
  g_coop_grid = max_workgroups;
  return 0;
}

int tfce_set_occupancy(int32_t wgs_per_cu) {
  if (wgs_per_cu < 1 || wgs_per_cu > 16) return set_error(TFCE_EINVAL, "occupancy must be in [1, 16]");
  g_wgs_per_cu = wgs_per_cu;
  return 0;
}

int tfce_adj_create_ex(const int32_t* row_ptr, const int32_t* col, int32_t V, int64_t nnz, float H,
                       float E, uint32_t flags, tfce_adj_t** out) {
  if (!out) return set_error(TFCE_EINVAL, "out is NULL");
  *out = nullptr;
  if (V <= 0 || V >= (1 << 24)) return set_error(TFCE_EINVAL, "V must be in [1, 2^24)");
  if (nnz < 0 || nnz >= ((int64_t)1 << 31)) return set_error(TFCE_EINVAL, "nnz must be in [0, 2^31)");
  if (!row_ptr || (nnz > 0 && !col)) return set_error(TFCE_EINVAL, "row_ptr/col is NULL");
  if (row_ptr[0] != 0 || row_ptr[V] != nnz) return set_error(TFCE_EADJ, "row_ptr[0] must be 0 and row_ptr[V] == nnz");
  for (int32_t v = 0; v < V; ++v)
    if (row_ptr[v + 1] < row_ptr[v]) return set_error(TFCE_EADJ, "row_ptr must be non-decreasing");
  for (int64_t e = 0; e < nnz; ++e)
    if (col[e] < 0 || col[e] >= V)
      return set_error(TFCE_EADJ, "adjacency index out of range at entry " + std::to_string(e));
  if ((flags & ~(uint32_t)(TFCE_ADJ_IMPLICIT_ZERO | TFCE_ADJ_ZERO_IN_SCALE)) != 0)
    return set_error(TFCE_EINVAL, "unknown adjacency flag");
  // symmetry: every (v,u) has (u,v); self loops and repeats are harmless
  bool symmetric = true;
  {
    std::vector<int64_t> fwd, bwd;
    fwd.reserve(nnz);
    bwd.reserve(nnz);
    for (int32_t v = 0; v < V; ++v)
      for (int32_t e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
        const int64_t u = col[e];
        if (u == v) continue;
        fwd.push_back((int64_t)v * V + u);
        bwd.push_back(u * V + v);
      }
    std::sort(fwd.begin(), fwd.end());
    fwd.erase(std::unique(fwd.begin(), fwd.end()), fwd.end());
    std::sort(bwd.begin(), bwd.end());
    bwd.erase(std::unique(bwd.begin(), bwd.end()), bwd.end());
    symmetric = (fwd == bwd);
  }
  tfce_adj* a = new tfce_adj();
  a->V = V;
  a->nnz = nnz;
  a->H = H;
  a->E = E;
  a->flags = (int32_t)flags | (symmetric ? kSurfSymmetric : 0);
  a->device = current_device();
  auto cleanup = [&](int code, const std::string& m) {
    if (a->d_row_ptr) (void)hipFree(a->d_row_ptr);
    if (a->d_col) (void)hipFree(a->d_col);
    if (a->d_ell) (void)hipFree(a->d_ell);
    if (a->d_nadj) (void)hipFree(a->d_nadj);
    if (a->d_adjw) (void)hipFree(a->d_adjw);
    if (a->d_adjid) (void)hipFree(a->d_adjid);
    if (a->d_sizepow) (void)hipFree(a->d_sizepow);
    if (a->d_surf) (void)hipFree(a->d_surf);
    delete a;
    return set_error(code, m);
  };
  hipError_t e;
  if ((e = hipMalloc(&a->d_row_ptr, sizeof(int32_t) * (V + 1))) != hipSuccess)
    return cleanup(-(int)e, "hipMalloc(row_ptr)");
  if ((e = hipMalloc(&a->d_col, sizeof(int32_t) * std::max<int64_t>(nnz, 1))) != hipSuccess)
    return cleanup(-(int)e, "hipMalloc(col)");
  if ((e = hipMalloc(&a->d_surf, sizeof(SurfDev))) != hipSuccess) return cleanup(-(int)e, "hipMalloc(surf)");
  if ((e = hipMemcpy(a->d_row_ptr, row_ptr, sizeof(int32_t) * (V + 1), hipMemcpyHostToDevice)) != hipSuccess)
    return cleanup(-(int)e, "upload row_ptr");
  if (nnz > 0 && (e = hipMemcpy(a->d_col, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice)) != hipSuccess)
    return cleanup(-(int)e, "upload col");
  // Fixed-width neighbour table: the sweep reads an element's neighbours from
  // one row instead of row_ptr then col (one random line fewer per element).
  {
    int32_t width = 0;
    for (int32_t v = 0; v < V; ++v) width = std::max(width, row_ptr[v + 1] - row_ptr[v]);
    width = (width + 3) & ~3;  // rows of whole 16-byte loads
    // rows wider than 16 entries (26-connected voxels) are padded to 32: one
    // 128-byte line per row, and 32-byte rows of adjacency bytes
    if (width > 16 && width <= kEllMaxWidth) width = kEllMaxWidth;
    if (width > 0 && width <= kEllMaxWidth) {
      std::vector<int32_t> ell((size_t)V * width, -1);
      for (int32_t v = 0; v < V; ++v)
        for (int32_t e2 = row_ptr[v]; e2 < row_ptr[v + 1]; ++e2) ell[(size_t)v * width + (e2 - row_ptr[v])] = col[e2];
      if ((e = hipMalloc(&a->d_ell, sizeof(int32_t) * ell.size())) != hipSuccess)
        return cleanup(-(int)e, "hipMalloc(ell)");
      if ((e = hipMemcpy(a->d_ell, ell.data(), sizeof(int32_t) * ell.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(-(int)e, "upload ell");
      a->ell_width = width;
      // Which neighbours of a row are adjacent to each other, within each
      // group of 8 consecutive entries (symmetric surfaces): two such
      // neighbours that were both active before a level are already in one
      // set, so the union step joins one of each adjacent group.
      if (symmetric && V < kPosMask) {
        std::vector<int32_t> srt(col, col + nnz);
        for (int32_t v = 0; v < V; ++v) std::sort(srt.begin() + row_ptr[v], srt.begin() + row_ptr[v + 1]);
        auto adjacent = [&](int32_t x, int32_t y) {
          return std::binary_search(srt.begin() + row_ptr[x], srt.begin() + row_ptr[x + 1], y);
        };
        int32_t max_deg = 0;
        for (int32_t v = 0; v < V; ++v) max_deg = std::max(max_deg, row_ptr[v + 1] - row_ptr[v]);
        if (max_deg <= kRingMaxDeg && V < kRingNone && width == 8) {
          // Rows of at most six neighbours: each row is put in ring order
          // (walk to an unvisited neighbour adjacent to the last one while
          // there is one) and the adjacency of ring-consecutive neighbours
          // kept as one link byte.  Adjacencies the walk does not place next
          // to each other are left out, which only keeps more candidates.
          // Each row is then packed into 16 bytes: six 18-bit indices, the
          // link byte from bit 108, the length from bit 114.
          std::vector<uint32_t> packed((size_t)V * 4, 0);
          for (int32_t v = 0; v < V; ++v) {
            int32_t* row = ell.data() + (size_t)v * width;
            const int deg = row_ptr[v + 1] - row_ptr[v];
            for (int j = 1; j < deg; ++j) {
              for (int k = j; k < deg; ++k)
                if (adjacent(row[j - 1], row[k])) {
                  std::swap(row[j], row[k]);
                  break;
                }
            }
            uint8_t m = 0;
            for (int j = 0; j + 1 < deg; ++j)
              if (adjacent(row[j], row[j + 1])) m |= (uint8_t)(1u << j);
            if (deg >= 3 && adjacent(row[deg - 1], row[0])) m |= (uint8_t)(1u << 5);
            uint64_t f[kRingMaxDeg];
            for (int j = 0; j < kRingMaxDeg; ++j) f[j] = j < deg ? (uint64_t)row[j] : (uint64_t)kRingNone;
            const uint64_t lo = f[0] | f[1] << 18 | f[2] << 36 | f[3] << 54;
            const uint64_t hi = f[3] >> 10 | f[4] << 8 | f[5] << 26 | (uint64_t)m << 44 | (uint64_t)deg << 50;
            packed[(size_t)v * 4 + 0] = (uint32_t)lo;
            packed[(size_t)v * 4 + 1] = (uint32_t)(lo >> 32);
            packed[(size_t)v * 4 + 2] = (uint32_t)hi;
            packed[(size_t)v * 4 + 3] = (uint32_t)(hi >> 32);
          }
          if ((e = hipMemcpy(a->d_ell, ell.data(), sizeof(int32_t) * ell.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup(-(int)e, "upload ring-ordered ell");
          const size_t pbytes = packed.size() * sizeof(uint32_t);
          if ((e = hipMalloc(&a->d_nadj, pbytes)) != hipSuccess) return cleanup(-(int)e, "hipMalloc(ring rows)");
          if ((e = hipMemcpy(a->d_nadj, packed.data(), pbytes, hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup(-(int)e, "upload ring rows");
          a->flags |= kSurfRing;
        } else {
          std::vector<uint8_t> nadj((size_t)V * width, 0);
          for (int32_t v = 0; v < V; ++v)
            for (int32_t j = 0; j < width; ++j) {
              const int32_t uj = ell[(size_t)v * width + j];
              if (uj < 0) continue;
              uint8_t m = 0;
              const int32_t c0 = j & ~7;
              for (int32_t k = c0; k < std::min(width, c0 + 8); ++k) {
                const int32_t uk = ell[(size_t)v * width + k];
                if (k != j && uk >= 0 && adjacent(uj, uk)) m |= (uint8_t)(1u << (k - c0));
              }
              nadj[(size_t)v * width + j] = m;
            }
          if ((e = hipMalloc(&a->d_nadj, nadj.size())) != hipSuccess) return cleanup(-(int)e, "hipMalloc(nadj)");
          if ((e = hipMemcpy(a->d_nadj, nadj.data(), nadj.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup(-(int)e, "upload nadj");
          if (width > kNposMaxWidth) {
            // wide rows: the adjacency among all of a row's neighbours, one
            // 32-bit mask per entry (bit k: adjacent to entry k), from which
            // the sweep groups an element's earlier neighbours exactly
            std::vector<uint32_t> adjw((size_t)V * width, 0u);
            std::vector<int32_t> mark(V, -1);  // entry of each neighbour in the current row
            for (int32_t v = 0; v < V; ++v) {
              const int32_t* row = ell.data() + (size_t)v * width;
              for (int32_t j = 0; j < width; ++j)
                if (row[j] >= 0) mark[row[j]] = j;
              for (int32_t j = 0; j < width; ++j) {
                const int32_t uj = row[j];
                if (uj < 0) continue;
                uint32_t m = 0u;
                for (int32_t e2 = row_ptr[uj]; e2 < row_ptr[uj + 1]; ++e2) {
                  const int32_t k = mark[col[e2]];
                  if (k >= 0 && k != j) m |= 1u << k;
                }
                adjw[(size_t)v * width + j] = m;
              }
              for (int32_t j = 0; j < width; ++j)
                if (row[j] >= 0) mark[row[j]] = -1;
            }
            // A grid's rows repeat (every interior voxel of a 26-connected
            // volume has the same neighbour offsets in the same order): keep
            // each distinct row once, in L2, and a 16-bit row id per element,
            // when there are few enough distinct rows
            std::vector<uint16_t> rid;
            std::vector<uint32_t> dict;
            {
              std::unordered_map<uint64_t, std::vector<int32_t>> seen;  // row hash -> distinct rows
              const size_t rb = (size_t)width * sizeof(uint32_t);
              rid.resize(V);
              for (int32_t v = 0; v < V && rid.size() == (size_t)V; ++v) {
                const uint32_t* row = adjw.data() + (size_t)v * width;
                uint64_t h = 1469598103934665603ull;
                for (int32_t j = 0; j < width; ++j) h = (h ^ row[j]) * 1099511628211ull;
                int32_t id = -1;
                for (int32_t d : seen[h])
                  if (std::memcmp(dict.data() + (size_t)d * width, row, rb) == 0) { id = d; break; }
                if (id < 0) {
                  id = (int32_t)(dict.size() / width);
                  if (id >= kAdjDictMaxRows) { rid.clear(); break; }
                  dict.insert(dict.end(), row, row + width);
                  seen[h].push_back(id);
                }
                rid[v] = (uint16_t)id;
              }
            }
            // (only when rows repeat: all-distinct rows would add an id load)
            const bool use_dict = rid.size() == (size_t)V && 2 * (dict.size() / width) <= (size_t)V;
            const std::vector<uint32_t>& wsrc = use_dict ? dict : adjw;
            const size_t wbytes = wsrc.size() * sizeof(uint32_t);
            if ((e = hipMalloc(&a->d_adjw, wbytes)) != hipSuccess) return cleanup(-(int)e, "hipMalloc(adjw)");
            if ((e = hipMemcpy(a->d_adjw, wsrc.data(), wbytes, hipMemcpyHostToDevice)) != hipSuccess)
              return cleanup(-(int)e, "upload adjw");
            if (use_dict) {
              const size_t ibytes = rid.size() * sizeof(uint16_t);
              if ((e = hipMalloc(&a->d_adjid, ibytes)) != hipSuccess) return cleanup(-(int)e, "hipMalloc(adjid)");
              if ((e = hipMemcpy(a->d_adjid, rid.data(), ibytes, hipMemcpyHostToDevice)) != hipSuccess)
                return cleanup(-(int)e, "upload adjid");
            }
          }
        }
      }
    }
  }
  if (E != 1.0f) {
    if ((e = hipMalloc(&a->d_sizepow, sizeof(double) * ((size_t)V + 1))) != hipSuccess)
      return cleanup(-(int)e, "hipMalloc(sizepow)");
    if ((e = launch_sizepow(a->d_sizepow, V, E, nullptr)) != hipSuccess || (e = hipDeviceSynchronize()) != hipSuccess)
      return cleanup(-(int)e, "sizepow table");
  }
  SurfDev sd;
  sd.row_ptr = a->d_row_ptr;
  sd.col = a->d_col;
  sd.ell = a->d_ell;
  sd.ell_width = a->ell_width;
  sd.nadj = a->d_nadj;
  sd.adjw = a->d_adjw;
  sd.adjid = a->d_adjid;
  sd.sizepow = a->d_sizepow;
  sd.offset = 0;
  sd.V = V;
  sd.flags = a->flags;
  sd.H = H;
  sd.E = E;
  if ((e = hipMemcpy(a->d_surf, &sd, sizeof(SurfDev), hipMemcpyHostToDevice)) != hipSuccess)
    return cleanup(-(int)e, "upload surface");
  *out = a;
  return 0;
}

// Reverse Cuthill-McKee order, component by component.  Each component starts
// at a pseudo-peripheral vertex (repeated BFS from a smallest-degree vertex of
// the last BFS level while the eccentricity grows); BFS visits unvisited
// neighbours in increasing degree, ties by index.  The graph is taken as
// given (out-lists), self loops ignored.
int tfce_locality_order(const int32_t* row_ptr, const int32_t* col, int32_t V, int32_t* new_of_old) {
  if (V <= 0 || !row_ptr || !new_of_old) return set_error(TFCE_EINVAL, "bad arguments");
  if (row_ptr[0] != 0) return set_error(TFCE_EADJ, "row_ptr[0] must be 0");
  for (int32_t v = 0; v < V; ++v)
    if (row_ptr[v + 1] < row_ptr[v]) return set_error(TFCE_EADJ, "row_ptr must be non-decreasing");
  if (row_ptr[V] > 0 && !col) return set_error(TFCE_EINVAL, "col is NULL");
  for (int64_t e = 0; e < row_ptr[V]; ++e)
    if (col[e] < 0 || col[e] >= V) return set_error(TFCE_EADJ, "adjacency index out of range");
  std::vector<int32_t> deg(V), order, mark(V, -1), level(V);
  for (int32_t v = 0; v < V; ++v) deg[v] = row_ptr[v + 1] - row_ptr[v];
  order.reserve(V);
  std::vector<char> done(V, 0);
  std::vector<int32_t> queue(V), nb;
  int stamp = 0;
  // BFS over the not-yet-ordered vertices from s; returns the vertices in visit
  // order (in queue[0, n)) and their BFS levels; sorted=true visits neighbours
  // by increasing degree
  auto bfs = [&](int32_t s, bool sorted) {
    ++stamp;
    int32_t head = 0, tail = 0;
    queue[tail++] = s;
    mark[s] = stamp;
    level[s] = 0;
    while (head < tail) {
      const int32_t v = queue[head++];
      nb.clear();
      for (int32_t e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
        const int32_t u = col[e];
        if (u != v && !done[u] && mark[u] != stamp) {
          mark[u] = stamp;
          level[u] = level[v] + 1;
          nb.push_back(u);
        }
      }
      if (sorted)
        std::sort(nb.begin(), nb.end(), [&](int32_t x, int32_t y) { return deg[x] != deg[y] ? deg[x] < deg[y] : x < y; });
      for (int32_t u : nb) queue[tail++] = u;
    }
    return tail;
  };
  for (int32_t s0 = 0; s0 < V; ++s0) {
    if (done[s0]) continue;
    // pseudo-peripheral start in s0's component
    int32_t s = s0;
    int32_t n = bfs(s, false);
    for (int32_t v = 0; v < n; ++v)
      if (deg[queue[v]] < deg[s] || (deg[queue[v]] == deg[s] && queue[v] < s)) s = queue[v];
    int ecc = -1;
    for (int iter = 0; iter < 8; ++iter) {
      n = bfs(s, false);
      const int e = level[queue[n - 1]];
      if (e <= ecc) break;
      ecc = e;
      int32_t best = -1;
      for (int32_t i = n - 1; i >= 0 && level[queue[i]] == e; --i)
        if (best < 0 || deg[queue[i]] < deg[best] || (deg[queue[i]] == deg[best] && queue[i] < best)) best = queue[i];
      s = best;
    }
    n = bfs(s, true);
    for (int32_t i = 0; i < n; ++i) {
      done[queue[i]] = 1;
      order.push_back(queue[i]);
    }
  }
  for (int32_t i = 0; i < V; ++i) new_of_old[order[i]] = V - 1 - i;
  return 0;
}

int tfce_adj_create(const int32_t* row_ptr, const int32_t* col, int32_t V, int64_t nnz, float H, float E,
                    tfce_adj_t** out) {
  return tfce_adj_create_ex(row_ptr, col, V, nnz, H, E, 0u, out);
}

int tfce_adj_destroy(tfce_adj_t* a) {
  if (!a) return 0;
  (void)hipFree(a->d_row_ptr);
  (void)hipFree(a->d_col);
  if (a->d_ell) (void)hipFree(a->d_ell);
  if (a->d_nadj) (void)hipFree(a->d_nadj);
  if (a->d_adjw) (void)hipFree(a->d_adjw);
  if (a->d_adjid) (void)hipFree(a->d_adjid);
  if (a->d_sizepow) (void)hipFree(a->d_sizepow);
  (void)hipFree(a->d_surf);
  delete a;
  return 0;
}

int tfce_adj_info(const tfce_adj_t* a, int32_t* V, int64_t* nnz, int32_t* symmetric) {
  if (!a) return set_error(TFCE_EINVAL, "adj is NULL");
  if (V) *V = a->V;
  if (nnz) *nnz = a->nnz;
  if (symmetric) *symmetric = (a->flags & kSurfSymmetric) ? 1 : 0;
  return 0;
}

// Words per element of the per-image row table the sweep keeps for a surface:
// neighbour positions (symmetric, fixed-width table of at most kNposMaxWidth),
// joined positions (wider symmetric tables with full row adjacency), 0 if none.
static int32_t np_width(const tfce_adj* a) {
  if (a->flags & kSurfRing) return kRingRowWords;
  if ((a->flags & kSurfSymmetric) && a->ell_width > 0 && a->ell_width <= kNposMaxWidth) return a->ell_width;
  return a->d_adjw ? kCandRowWords : 0;
}

// Whether the LDS-resident sweep takes a surface: symmetric, with a neighbour
// table, at most kLdsMaxV elements.
static bool lds_eligible(const tfce_adj* a) {
  return (a->flags & kSurfSymmetric) && a->d_ell != nullptr && a->V <= kLdsMaxV && g_lds_sweep;
}

static int run_batch(const SurfDev* d_surfs, int32_t n_surf, int32_t v_max, int32_t np_w, const float* d_maps,
                     int64_t map_stride, int P, int sign_mask, const float* d_T0,
                     const float* d_vdensity, bool full, float* d_enhn, float min_t, float dh,
                     int32_t labels_level, int32_t* d_labels, float* d_scaled, float* d_raw,
                     float* d_own, int32_t* d_status, hipStream_t stream, int out_mul = 1,
                     int out_add = 0, uint32_t batch_flags = 0, const ExtScratch* ext = nullptr,
                     bool all_ring = false, int32_t n_lds_surf = 0, bool rest_ring = false) {
  if (P <= 0) return 0;
  if (sign_mask < 1 || sign_mask > 3) return set_error(TFCE_EINVAL, "sign_mask must be 1, 2 or 3");
  const int n_signs = (sign_mask == 3) ? 2 : 1;
  const int64_t n_items = (int64_t)n_surf * P * n_signs;
  const int64_t slot_words = tfce_slot_words(v_max, np_w);
  const size_t slot_bytes = (size_t)slot_words * sizeof(int32_t);
  // resident workgroups, each with its own slot; the slots stay within a
  // quarter of the device's memory (fewer workgroups for very large surfaces)
  const int64_t mem_cap = (int64_t)(device_memory(current_device()) / 4 / slot_bytes);
  // the trailing n_lds_surf surfaces (the smallest) run in the LDS-resident
  // sweep, one 1,024-thread workgroup per compute unit, after the others
  const int64_t n_lds_items = full ? 0 : (int64_t)std::min(n_lds_surf, n_surf) * P * n_signs;
  const int64_t n_global = n_items - n_lds_items;
  int grid = (int)std::max<int64_t>(1, std::min<int64_t>(tfce_grid(std::max<int64_t>(n_global, 1)), mem_cap));
  const size_t prep_bytes = (size_t)n_items * kPrepWords * sizeof(int32_t);
  // raw_max is mandatory for the kernel: park it in scratch when not wanted
  const size_t raw_bytes = (size_t)n_items * sizeof(float);
  const size_t t0_bytes = (batch_flags & TFCE_BATCH_SHARED_THRESHOLDS) ? (size_t)n_items * sizeof(float) : 0;
  char* slots = nullptr;  // grid slots
  char* rest = nullptr;   // prep records, raw maxima, shared thresholds
  int32_t* queue = nullptr;
  if (ext) {
    // caller-given scratch: [queue 256 B][prep][raw, 256-aligned][T0][slots, 256-aligned]
    const size_t fixed =
        256 + ((prep_bytes + 255) / 256) * 256 + ((raw_bytes + 255) / 256) * 256 + ((t0_bytes + 255) / 256) * 256;
    if (fixed + slot_bytes > ext->bytes)
      return set_error(TFCE_EINVAL, "the batch needs " + std::to_string(fixed + slot_bytes) +
                                        " bytes of scratch (records of every item and one slot of the largest "
                                        "surface); the no-allocation scratch holds " + std::to_string(ext->bytes));
    grid = (int)std::min<int64_t>(grid, (int64_t)((ext->bytes - fixed) / slot_bytes));
    char* base = reinterpret_cast<char*>(ext->base);
    queue = reinterpret_cast<int32_t*>(base);
    rest = base + 256;
    slots = base + fixed;
  } else {
    Workspace* ws = nullptr;
    const size_t slots_bytes = (size_t)grid * slot_bytes;
    int rc = workspace_get(stream, slots_bytes + prep_bytes + raw_bytes + t0_bytes + 512, &ws);
    if (rc) return rc;
    slots = reinterpret_cast<char*>(ws->ptr);
    rest = slots + slots_bytes;
    queue = ws->queue;
  }
  TfceBatchArgs a{};
  a.surfs = d_surfs;
  a.n_surf = n_surf;
  a.n_maps = P;
  a.n_signs = n_signs;
  a.sign0 = (sign_mask == 2) ? -1.0f : 1.0f;
  a.maps = d_maps;
  a.map_stride = map_stride;
  a.ext_T0 = d_T0;
  a.vdensity = d_vdensity;
  a.min_t = min_t;
  a.dh = dh;
  a.full_scale = 0;
  a.scaled_max = d_scaled;
  a.out_mul = out_mul;
  a.out_add = out_add;
  a.own_max = d_own;
  a.status = d_status;
  a.enhn = d_enhn;
  a.labels_level = labels_level;
  a.labels = d_labels;
  a.scratch = reinterpret_cast<int32_t*>(slots);
  a.slot_stride = slot_words;
  a.v_max = v_max;
  a.np_w = np_w;
  a.queue = queue;
  a.n_items = n_items;
  a.prep = reinterpret_cast<int32_t*>(rest);
  a.raw_max = d_raw ? d_raw : reinterpret_cast<float*>(rest + prep_bytes);
  if (batch_flags & TFCE_BATCH_SHARED_THRESHOLDS) {
    float* T0 = reinterpret_cast<float*>(rest + prep_bytes + ((raw_bytes + 255) / 256) * 256);
    // (inside the prep + raw + T0 extent of either layout: the stream's
    // scratch adds 512 bytes, the caller's layout rounds each part to 256)
    hipError_t e0 = launch_shared_T0(a, T0, (batch_flags & TFCE_BATCH_ZERO_IN_IMAGE) ? 1 : 0, stream);
    if (e0 != hipSuccess) return set_error(-(int)e0, "shared threshold launch");
    a.ext_T0 = T0;
  }
  // the ring-only sweep kernel when every surface the global sweep takes is
  // ring-packed: all of them, or those left after the LDS-resident ones
  const bool ring_only = n_lds_items > 0 ? rest_ring : all_ring;
  // both signs of a map in one workgroup: MAX mode over ring-packed surfaces,
  // thresholds >= 0 (the two images' active sets are then disjoint); two
  // 512-thread workgroups per compute unit, within the slots allocated above
  int pair_grid = 0;
  if (g_pair_sweep && !full && n_signs == 2 && ring_only && min_t >= 0.0f && n_global >= 2)
    pair_grid = (int)std::max<int64_t>(1, std::min<int64_t>({n_global / 2, (int64_t)grid,
                                                              2 * (int64_t)cu_count(current_device())}));
  hipError_t e = launch_tfce_batch(a, full, ring_only, grid, n_global, stream, pair_grid);
  if (e != hipSuccess) return set_error(-(int)e, std::string("tfce launch: ") + hipGetErrorString(e));
  if (n_lds_items > 0) {
    // one workgroup per compute unit, each taking both signs of a map at a time
    const int lgrid = (int)std::max<int64_t>(1, std::min<int64_t>(n_lds_items / n_signs, cu_count(current_device())));
    e = launch_tfce_lds(a, n_global, lgrid, stream);
    if (e != hipSuccess) return set_error(-(int)e, std::string("tfce LDS launch: ") + hipGetErrorString(e));
  }
  return 0;
}

}  // extern "C"

namespace tfcehip {

int batch_max_in(tfce_adj_t* adj, const float* d_maps, int P, int sign_mask, const float* d_vdensity,
                 float* d_max, const ExtScratch* ext, hipStream_t stream) {
  if (!adj || !d_maps || !d_max) return set_error(TFCE_EINVAL, "NULL argument");
  if (P <= 0) return 0;
  if (sign_mask < 1 || sign_mask > 3) return set_error(TFCE_EINVAL, "sign_mask must be 1, 2 or 3");
  if (sign_mask != 3) {
    // the unrequested sign reads 0; the requested one lands at 2p + (sign_mask == 2)
    hipError_t e = launch_fill_f32(d_max, (int64_t)P * 2, 0.0f, stream);
    if (e != hipSuccess) return set_error(-(int)e, "fill max");
  }
  return run_batch(adj->d_surf, 1, adj->V, np_width(adj), d_maps, adj->V, P, sign_mask, nullptr, d_vdensity,
                   d_vdensity != nullptr, nullptr, 0.0f, 0.0f, -1, nullptr, d_max, nullptr, nullptr, nullptr,
                   stream, sign_mask == 3 ? 1 : 2, sign_mask == 2 ? 1 : 0, 0u, ext, (adj->flags & kSurfRing) != 0,
                   lds_eligible(adj) ? 1 : 0);
}

}  // namespace tfcehip

extern "C" {

}  // extern "C"

// Whether the current device takes cooperative launches (asked once per device).
static bool coop_supported() {
  static std::mutex mu;
  static std::map<int, bool> known;
  const int dev = current_device();
  std::lock_guard<std::mutex> g(mu);
  auto it = known.find(dev);
  if (it != known.end()) return it->second;
  int v = 0;
  const bool ok = hipDeviceGetAttribute(&v, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess && v != 0;
  known[dev] = ok;
  return ok;
}

// One image over the whole GPU: its record, then the cooperative level sweep.
static int run_coop(tfce_adj_t* adj, const float* d_image, float* d_enhn, float min_t, float dh,
                    int32_t* d_status, hipStream_t stream) {
  const int64_t V = adj->V;
  const size_t rec_bytes = ((size_t)kPrepWords * sizeof(int32_t) + 255) / 256 * 256;
  const size_t arr_bytes = ((size_t)V * sizeof(int32_t) + 255) / 256 * 256;
  const size_t ctl_bytes = (size_t)kCoopCtlWords * sizeof(int32_t);
  Workspace* ws = nullptr;
  int rc = workspace_get(stream, rec_bytes + 256 + 8 * arr_bytes + ctl_bytes + 256, &ws);
  if (rc) return rc;
  char* base = reinterpret_cast<char*>(ws->ptr);
  int32_t* rec = reinterpret_cast<int32_t*>(base);
  float* raw = reinterpret_cast<float*>(base + rec_bytes);
  char* arr = base + rec_bytes + 256;
  TfceBatchArgs a{};
  a.surfs = adj->d_surf;
  a.n_surf = 1;
  a.n_maps = 1;
  a.n_signs = 1;
  a.sign0 = 1.0f;
  a.maps = d_image;
  a.map_stride = V;
  a.min_t = min_t;
  a.dh = dh;
  a.raw_max = raw;
  a.queue = ws->queue;
  a.n_items = 1;
  a.prep = rec;
  a.v_max = (int32_t)V;
  hipError_t e = launch_tfce_batch(a, true, false, 1, 0, stream);
  if (e != hipSuccess) return set_error(-(int)e, std::string("tfce record launch: ") + hipGetErrorString(e));
  TfceCoopArgs c{};
  c.row_ptr = adj->d_row_ptr;
  c.col = adj->d_col;
  c.sizepow = adj->d_sizepow;
  c.E = adj->E;
  c.V = (int32_t)V;
  c.x = d_image;
  c.enhn = d_enhn;
  c.rec = rec;
  int32_t* w[8];
  for (int q = 0; q < 8; ++q) w[q] = reinterpret_cast<int32_t*>(arr + q * arr_bytes);
  c.lev = w[0];
  c.order = w[1];
  c.parent = w[2];
  c.size = w[3];
  c.roots[0] = w[4];
  c.roots[1] = w[5];
  c.rootc = w[6];
  c.acc = reinterpret_cast<float*>(w[7]);
  c.ctl = reinterpret_cast<int32_t*>(arr + 8 * arr_bytes);
  // the grid barrier's counter and generation (the ctl region's last words)
  e = launch_fill_f32(reinterpret_cast<float*>(c.ctl + kCoopBarWord), 2, 0.0f, stream);
  if (e != hipSuccess) return set_error(-(int)e, "clear barrier words");
  c.status = d_status;
  c.grid_limit = g_coop_grid;
  e = launch_tfce_coop(c, stream);
  if (e != hipSuccess) return set_error(-(int)e, std::string("tfce cooperative launch: ") + hipGetErrorString(e));
  return 0;
}

extern "C" {

int tfce_run_ex(tfce_adj_t* adj, const float* d_image, float* d_enhn, float min_t, float dh, int accumulate,
                int32_t labels_level, int32_t* d_labels, int32_t* d_status, hipStream_t stream) {
  if (!adj || !d_image || !d_enhn) return set_error(TFCE_EINVAL, "NULL argument");
  if (dh < 0.0f) return set_error(TFCE_EINVAL, "dh must be >= 0");
  if (!accumulate) {
    hipError_t e = launch_fill_f32(d_enhn, adj->V, 0.0f, stream);
    if (e != hipSuccess) return set_error(-(int)e, "fill enhn");
  }
  // a large symmetric surface runs over the whole GPU (the labels debug
  // output exists only in the one-workgroup sweep) where the device takes
  // cooperative launches
  if ((adj->flags & kSurfSymmetric) && labels_level < 0 &&
      (g_coop_mode == 1 || (g_coop_mode == -1 && adj->V >= kCoopMinV)) && coop_supported())
    return run_coop(adj, d_image, d_enhn, min_t, dh, d_status, stream);
  return run_batch(adj->d_surf, 1, adj->V, np_width(adj), d_image, adj->V, 1, 1, nullptr, nullptr, true, d_enhn, min_t,
                   dh, labels_level, d_labels, nullptr, nullptr, nullptr, d_status, stream, 1, 0, 0u, nullptr,
                   (adj->flags & kSurfRing) != 0);
}

int tfce_run(tfce_adj_t* adj, const float* d_image, float* d_enhn, float dh, int accumulate,
             hipStream_t stream) {
  return tfce_run_ex(adj, d_image, d_enhn, 0.0f, dh, accumulate, -1, nullptr, nullptr, stream);
}

int tfce_batch_max(tfce_adj_t* adj, const float* d_maps, int P, int sign_mask, const float* d_vdensity,
                   float* d_max, hipStream_t stream) {
  return tfcehip::batch_max_in(adj, d_maps, P, sign_mask, d_vdensity, d_max, nullptr, stream);
}

}  // extern "C"

// Number of trailing surfaces (smallest first from the end) the LDS-resident
// sweep takes, and whether the others are all ring-packed.
static int32_t lds_tail(const tfce_surfset* s) {
  if (!g_lds_sweep) return 0;
  int32_t k = 0;
  while (k < s->n && s->lds_ok[s->n - 1 - k]) ++k;
  return k;
}

static bool big_all_ring(const tfce_surfset* s) {
  const int32_t n_big = s->n - lds_tail(s);
  for (int32_t k = 0; k < n_big; ++k)
    if (!(s->host[k].flags & kSurfRing)) return false;
  return true;
}

extern "C" {

int tfce_surfset_create(const tfce_adj_t* const* adjs, int32_t n, const int64_t* offsets, tfce_surfset_t** out) {
  if (!out || !adjs || !offsets || n <= 0) return set_error(TFCE_EINVAL, "bad surface set arguments");
  *out = nullptr;
  tfce_surfset* s = new tfce_surfset();
  s->n = n;
  s->device = current_device();
  std::vector<int32_t> idx(n);
  for (int32_t i = 0; i < n; ++i) {
    if (!adjs[i]) { delete s; return set_error(TFCE_EINVAL, "NULL adjacency in set"); }
    idx[i] = i;
  }
  // launch order: the surfaces the LDS-resident sweep can take last (they form
  // its item range), largest first within each group (work balance of the
  // persistent queues)
  auto small = [&](int32_t i) {
    const tfce_adj* a = adjs[i];
    return (a->flags & kSurfSymmetric) && a->d_ell != nullptr && a->V <= kLdsMaxV;
  };
  std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
    if (small(a) != small(b)) return !small(a);
    return adjs[a]->V > adjs[b]->V;
  });
  for (int32_t k = 0; k < n; ++k) {
    const tfce_adj* a = adjs[idx[k]];
    SurfDev sd;
    sd.row_ptr = a->d_row_ptr;
    sd.col = a->d_col;
    sd.ell = a->d_ell;
    sd.ell_width = a->ell_width;
    sd.nadj = a->d_nadj;
    sd.adjw = a->d_adjw;
    sd.adjid = a->d_adjid;
    sd.sizepow = a->d_sizepow;
    sd.offset = offsets[idx[k]];
    sd.V = a->V;
    sd.flags = a->flags;
    sd.H = a->H;
    sd.E = a->E;
    s->host.push_back(sd);
    s->user_index.push_back(idx[k]);
    s->v_max = std::max(s->v_max, a->V);
    s->np_w = std::max(s->np_w, np_width(a));
    s->all_ring = s->all_ring && (a->flags & kSurfRing) != 0;
    s->lds_ok.push_back(small(idx[k]) ? 1 : 0);
  }
  hipError_t e = hipMalloc(&s->d_surfs, sizeof(SurfDev) * n);
  if (e == hipSuccess) e = hipMemcpy(s->d_surfs, s->host.data(), sizeof(SurfDev) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (s->d_surfs) (void)hipFree(s->d_surfs);
    delete s;
    return set_error(-(int)e, "upload surface set");
  }
  *out = s;
  return 0;
}

int tfce_surfset_destroy(tfce_surfset_t* s) {
  if (!s) return 0;
  (void)hipFree(s->d_surfs);
  delete s;
  return 0;
}

int tfce_batch_ex(const tfce_surfset_t* s, const float* d_maps, int64_t map_stride, int P, int sign_mask,
                  uint32_t flags, const float* d_vdensity, float* d_scaled, float* d_raw, float* d_own,
                  int32_t* d_status, hipStream_t stream) {
  if (!s || !d_maps || !d_scaled) return set_error(TFCE_EINVAL, "NULL argument");
  if (flags & ~(uint32_t)(TFCE_BATCH_SHARED_THRESHOLDS | TFCE_BATCH_ZERO_IN_IMAGE))
    return set_error(TFCE_EINVAL, "unknown batch flag");
  return run_batch(s->d_surfs, s->n, s->v_max, s->np_w, d_maps, map_stride, P, sign_mask, nullptr, d_vdensity,
                   d_vdensity != nullptr, nullptr, 0.0f, 0.0f, -1, nullptr, d_scaled, d_raw, d_own, d_status,
                   stream, 1, 0, flags, nullptr, s->all_ring, lds_tail(s), big_all_ring(s));
}

int tfce_surfset_order(const tfce_surfset_t* s, int32_t* user_index) {
  if (!s || !user_index) return set_error(TFCE_EINVAL, "NULL argument");
  for (int32_t k = 0; k < s->n; ++k) user_index[k] = s->user_index[k];
  return 0;
}

int tfce_batch(const tfce_surfset_t* s, const float* d_maps, int64_t map_stride, int P, int sign_mask,
               const float* d_T0, const float* d_vdensity, float* d_scaled, float* d_raw, float* d_own,
               int32_t* d_status, hipStream_t stream) {
  if (!s || !d_maps || !d_scaled) return set_error(TFCE_EINVAL, "NULL argument");
  return run_batch(s->d_surfs, s->n, s->v_max, s->np_w, d_maps, map_stride, P, sign_mask, d_T0, d_vdensity,
                   d_vdensity != nullptr, nullptr, 0.0f, 0.0f, -1, nullptr, d_scaled, d_raw, d_own, d_status,
                   stream, 1, 0, 0u, nullptr, s->all_ring, lds_tail(s), big_all_ring(s));
}

}  // extern "C"
