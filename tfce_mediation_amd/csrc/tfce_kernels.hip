// Batched TFCE on CDNA4 (gfx950).
//
// Semantics follow tfce_mediation/lib/fast_tfce.hpp:11-95 (the operator behind
// CreateAdjSet.run, tfce.pyx:24-45): thresholds T_0 = max, T_{i+1} = fl(T_i - dT)
// while T_i >= minT, dT = fl(max/100); a vertex is active at T iff x > T; each
// member of a component of size s gains (float)(s^E * fl(T^H)) per level, added
// in float in descending-T order.
//
// Design (not the reference's): one work item = one (surface, map, sign) image.
//   tfce_prepare_kernel - one workgroup per item, all items at once: the
//     image maximum and the float threshold / T^H tables of the item.
//   tfce_batch_kernel   - a persistent grid pulls items from an atomic queue;
//     each workgroup owns one scratch slot and runs the item's sweep:
//     1. bucket the vertices by activation level (counting sort, LDS
//        histogram); compact id = position in that order, so the active set
//        at level l is the id prefix [0, end_l) and "is u active" is
//        `pos[u] < end_l`; for symmetric surfaces with a fixed-width
//        neighbour table, one row of neighbour positions per active element
//        (a joined neighbour is then simply one with a smaller position), or
//        for the wide voxel tables one row of the positions to join (the
//        earlier neighbours, one per adjacent group of older ones);
//     2. per level: hook the newly active vertices into a lock-free union-find
//        (roots of an element's joined neighbours found as interleaved
//        chains, every other root hooked under the smallest by CAS, larger
//        under smaller), merge component sizes into the surviving roots with
//        atomics, then
//          MAX mode : add the level increment to one value per root; a merged
//                     root keeps the larger value (atomicMax on float bits); the
//                     largest TFCE of a component is held by its earliest vertex,
//                     so max over roots == max over vertices, bit for bit;
//          FULL mode: add the increment of its root to every active vertex.
//   tfce_pair_kernel    - the MAX sweep of ring-packed surfaces with both signs
//     of a map in one workgroup (the two images' active sets are disjoint and
//     share one slot), one pass over the map and the vertices for the two.
//   tfce_lds_kernel     - the MAX sweep of small surfaces with the whole
//     union-find of a map's two images in LDS.
//   tfce_coop_kernel    - one image over the whole GPU (cooperative launch).
// Every level touches only the new vertices and the current roots (MAX) or the
// active prefix (FULL).  The sweep kernel carries no transcendental math (T^H
// comes from the prepare kernel, s^E from a per-adjacency table) so it keeps
// few registers and many waves resident: it is latency bound.
#include "tfce_internal.h"

namespace tfcehip {

namespace {

__device__ __forceinline__ int ld(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void st(int32_t* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int uf_find(int32_t* parent, int x) {
  int p = ld(parent + x);
  while (p != x) {
    int gp = ld(parent + p);
    if (gp != p) st(parent + x, gp);  // path halving; gp is an ancestor of x
    x = p;
    p = gp;
  }
  return x;
}

// Join the sets of a and b, starting from a (an element or any ancestor of it);
// returns the root the two sets had when joined, a good start for a's next
// union.  The larger root is hooked under the smaller by CAS; a failed CAS
// (the root was hooked meanwhile) retries from the new roots.
// A root of this level's elements (id >= b0) that gets hooked gives its own
// unit of size to its new parent at once; whatever it had received itself is
// passed on by the merge step.
__device__ __forceinline__ int uf_link(int32_t* parent, int32_t* size, int a, int b, int b0) {
  while (true) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return a;
    if (a < b) { int t = a; a = b; b = t; }
    int expected = a;
    if (__hip_atomic_compare_exchange_strong(parent + a, &expected, b, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
      if (a >= b0) __hip_atomic_fetch_add(size + b, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      return b;
    }
  }
}

// (float)(s^E * HH) for one component: E == 1 is one correctly rounded float
// product (exact in double, as fast_tfce.hpp:77 computes it); otherwise the
// per-adjacency table holds (double)s^E.
__device__ __forceinline__ float tfce_inc(int size, float HH, float E, const double* sizepow) {
  if (E == 1.0f) return __fmul_rn((float)size, HH);
  return __double2float_rn(sizepow[size] * (double)HH);
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// max over the workgroup; NaN inputs are dropped by fmaxf
__device__ float block_max(float v, float* s_red) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) s_red[w] = v;
  __syncthreads();
  float r = s_red[0];
  for (int i = 1; i < kTfceThreads / 64; ++i) r = fmaxf(r, s_red[i]);
  __syncthreads();
  return r;
}

// Append `c` to list when `take` (wave-aggregated LDS counter).  Every lane
// of the wave must call it (it ballots).
__device__ __forceinline__ void wave_append(bool take, int c, int32_t* list, int* s_count) {
  unsigned long long m = __ballot(take);
  if (m == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(s_count, __popcll(m));
  base = __shfl(base, leader, 64);
  if (take) list[base + __popcll(m & ((1ull << lane) - 1ull))] = c;
}

constexpr int kIlp = 4;  // independent find chains per thread in the merge phases

// Roots of up to N elements at once (x[j] < 0: none), the loads of the
// chains interleaved.  SPLIT: path splitting, a visited node is re-pointed to
// its grandparent.  The union step with neighbour-position rows only reads
// (its writes cost more than the shorter chains save there; the merge step
// compresses); with the wide voxel tables splitting pays.
template <int N, bool SPLIT = true>
__device__ __forceinline__ void find_many(int32_t* parent, int (&x)[N]) {
  int prev[N];
  bool go[N];
  bool any = false;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    go[j] = x[j] >= 0;
    prev[j] = -1;
    any |= go[j];
  }
  while (any) {
    int p[N];
#pragma unroll
    for (int j = 0; j < N; ++j) p[j] = go[j] ? ld(parent + x[j]) : 0;
    any = false;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if (!go[j]) continue;
      if (p[j] == x[j]) {
        go[j] = false;
      } else {
        if (SPLIT && prev[j] >= 0) st(parent + prev[j], p[j]);
        prev[j] = x[j];
        x[j] = p[j];
        any = true;
      }
    }
  }
}

// Roots of the live ones of kIlp elements (find_many on them).
__device__ __forceinline__ void find_batch(int32_t* parent, int (&x)[kIlp], const bool (&live)[kIlp]) {
  int y[kIlp];
#pragma unroll
  for (int j = 0; j < kIlp; ++j) y[j] = live[j] ? x[j] : -1;
  find_many<kIlp>(parent, y);
#pragma unroll
  for (int j = 0; j < kIlp; ++j)
    if (live[j]) x[j] = y[j];
}

// Adjacency bytes of entries e..e+7 of vertex v's table row (zeros past the
// row's end), as two words.
__device__ __forceinline__ void chunk_adj(const uint8_t* nadj, int v, int W, int e, unsigned& wlo, unsigned& whi) {
  if (!nadj) return;
  const uint32_t* p = reinterpret_cast<const uint32_t*>(nadj + (size_t)v * W + e);
  wlo = p[0];
  whi = (e + 4 < W) ? p[1] : 0u;
}

template <bool SPLIT>
__device__ __forceinline__ int join_roots(int32_t* parent, int32_t* size, int (&r)[9], int ri, int b0);

// Neighbour-position entry of a surface with a local adjacency table.
__device__ __forceinline__ int np_encode(int q, const uint8_t* nadj_row, int j) {
  return (q == kNever ? kPosNone : q) | ((int)nadj_row[j] << 24);
}

// Candidates of element i from its row of encoded entries (width <= 8): every
// joined neighbour of this level, and one neighbour per group of mutually
// adjacent neighbours that were active before it (such a group is one set
// already: its edges were joined at earlier levels).
__device__ __forceinline__ void prune8(const int (&p)[8], const unsigned (&adj)[8], int i, int b0,
                                       int (&r)[9]) {
  unsigned old_m = 0, keep = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (p[j] < i) {
      if (p[j] < b0) old_m |= 1u << j;
      else keep |= 1u << j;
    }
  }
  unsigned rem = old_m;
  while (rem) {
    const unsigned seed = rem & (0u - rem);
    unsigned comp = seed, prev;
    do {
      prev = comp;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (comp & (1u << j)) comp |= adj[j] & old_m;
    } while (comp != prev);
    keep |= seed;
    rem &= ~comp;
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = (keep >> j) & 1u ? p[j] : -1;
}

// Ring-packed row of one element: six 18-bit fields from bit 0 (neighbour
// indices in the table, positions in the per-image rows; kRingNone: none or
// inactive), then the 9-bit tail: link byte (bits 108-113), row length
// (bits 114-116).
__device__ __forceinline__ int4 ring_pack(const int (&q)[kRingMaxDeg], unsigned tail) {
  auto c = [](int v) -> uint64_t { return (uint64_t)(uint32_t)(v == kNever ? kRingNone : v); };
  const uint64_t lo = c(q[0]) | c(q[1]) << 18 | c(q[2]) << 36 | c(q[3]) << 54;
  const uint64_t hi = c(q[3]) >> 10 | c(q[4]) << 8 | c(q[5]) << 26 | (uint64_t)tail << 44;
  return make_int4((int)(uint32_t)lo, (int)(uint32_t)(lo >> 32), (int)(uint32_t)hi, (int)(uint32_t)(hi >> 32));
}

__device__ __forceinline__ void ring_fields(const int4 w, int (&p)[kRingMaxDeg], unsigned& tail) {
  const uint64_t lo = (uint64_t)(uint32_t)w.x | (uint64_t)(uint32_t)w.y << 32;
  const uint64_t hi = (uint64_t)(uint32_t)w.z | (uint64_t)(uint32_t)w.w << 32;
  const uint64_t M = kRingNone;
  p[0] = (int)(lo & M);
  p[1] = (int)((lo >> 18) & M);
  p[2] = (int)((lo >> 36) & M);
  p[3] = (int)(((lo >> 54) | (hi << 10)) & M);
  p[4] = (int)((hi >> 8) & M);
  p[5] = (int)((hi >> 26) & M);
  tail = (unsigned)(hi >> 44) & 0x1ffu;
}

// Positions (kRingNone: inactive or no entry) and the adjacency masks among
// them of a ring-packed row.
__device__ __forceinline__ void ring_unpack(const int4 w, int (&p)[8], unsigned (&adj)[8]) {
  int f[kRingMaxDeg];
  unsigned tail;
  ring_fields(w, f, tail);
#pragma unroll
  for (int j = 0; j < kRingMaxDeg; ++j) p[j] = f[j];
  p[6] = p[7] = kRingNone;
  const unsigned link = tail & 63u;
  const int deg = (int)(tail >> 6);
#pragma unroll
  for (int j = 0; j < 8; ++j) adj[j] = 0;
#pragma unroll
  for (int j = 0; j < 5; ++j)
    if ((link >> j) & 1u) {
      adj[j] |= 2u << j;
      adj[j + 1] |= 1u << j;
    }
  if (link & 32u) {
#pragma unroll
    for (int j = 2; j < kRingMaxDeg; ++j)
      if (j == deg - 1) {
        adj[j] |= 1u;
        adj[0] |= 1u << j;
      }
  }
}

__device__ __forceinline__ void prune_row(const int (&e)[8], int i, int b0, int (&r)[9]) {
  int p[8];
  unsigned adj[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    p[j] = e[j] & kPosMask;
    adj[j] = (unsigned)e[j] >> 24;
  }
  prune8(p, adj, i, b0, r);
}

// 3b for one element i (vertex v) and up to 8 of its neighbours u[] (-1: none)
// of a table or CSR row without adjacency bytes: gather their positions, find
// the roots of the joined ones and of i's set (ri, an ancestor) as interleaved
// chains, then hook every other distinct root under the smallest with
// independent CASes; a failed CAS (the root was hooked meanwhile) falls back
// to uf_link.  Returns an ancestor of i's set.
__device__ __forceinline__ int join_chunk(int32_t* parent, int32_t* size, const int32_t* pos, const float* x,
                                          const int (&u)[8],
                                          int i, int v, int ri, int b0, int b1, bool symmetric, float sign,
                                          float xv) {
  int q[8], r[9];
#pragma unroll
  for (int j = 0; j < 8; ++j) q[j] = u[j] >= 0 ? pos[u[j]] : kNever;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    bool take = q[j] < b1 && q[j] != i;
    if (take && q[j] >= b0) {
      if (symmetric) {
        take = q[j] < i;  // a same-level edge is joined once
      } else {
        // listed from v only: the reference uses it iff u was activated
        // first (descending value, ties by index)
        const float xu = sign * x[u[j]];
        take = xu > xv || (xu == xv && u[j] < v);
      }
    }
    r[j] = take ? q[j] : -1;
  }
  return join_roots<true>(parent, size, r, ri, b0);
}

// Joins element i (vertex row `row` of a symmetric table of width W) with
// every neighbour of a smaller position, eight entries at a time, without
// pruning: the path of elements whose joined-position row overflowed.
__device__ int join_table_row(int32_t* parent, int32_t* size, const int32_t* pos, const int32_t* row, int W, int i,
                              int b0) {
  int ri = i;
  for (int e = 0; e < W; e += 8) {
    const int4 lo = reinterpret_cast<const int4*>(row)[e >> 2];
    const int4 hi = (e + 4 < W) ? reinterpret_cast<const int4*>(row)[(e >> 2) + 1] : make_int4(-1, -1, -1, -1);
    const int u[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    int r[9];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int q = u[j] >= 0 ? pos[u[j]] : kNever;
      r[j] = q < i ? q : -1;
    }
    ri = join_roots<true>(parent, size, r, ri, b0);
    if (u[7] < 0) break;  // -1 padding: the row has ended
  }
  return ri;
}

// Compare-exchange for a descending sorting network.
__device__ __forceinline__ void cx_desc(int& a, int& b) {
  const int hi = max(a, b), lo = min(a, b);
  a = hi;
  b = lo;
}

// r[0..7] sorted descending (Batcher's 19-comparator network).
__device__ __forceinline__ void sort8_desc(int (&r)[9]) {
  cx_desc(r[0], r[1]); cx_desc(r[2], r[3]); cx_desc(r[4], r[5]); cx_desc(r[6], r[7]);
  cx_desc(r[0], r[2]); cx_desc(r[1], r[3]); cx_desc(r[4], r[6]); cx_desc(r[5], r[7]);
  cx_desc(r[1], r[2]); cx_desc(r[5], r[6]); cx_desc(r[0], r[4]); cx_desc(r[3], r[7]);
  cx_desc(r[1], r[5]); cx_desc(r[2], r[6]); cx_desc(r[1], r[4]); cx_desc(r[3], r[6]);
  cx_desc(r[2], r[4]); cx_desc(r[3], r[5]); cx_desc(r[3], r[4]);
}

// The root-joining half of join_chunk: r[0..7] are positions to join with
// i's set (ri, an ancestor of it), -1 for none.  The positions are sorted
// descending first (Batcher's 19-comparator network), so the empty slots are
// the last ones and a slot that no lane of the wave uses issues no memory
// instruction in the chains and CASes below.
template <bool SPLIT>
__device__ __forceinline__ int join_roots(int32_t* parent, int32_t* size, int (&r)[9], int ri, int b0) {
  sort8_desc(r);
  // i's own set enters as ri without a lookup: if ri is no longer a root its
  // CAS below fails and uf_link takes over, and if it is the smallest, the
  // other roots hooked under it still land in its set under a smaller id
  int rn[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) rn[j] = r[j];
  find_many<8, SPLIT>(parent, rn);
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = rn[j];
  r[8] = ri;
  int tgt = r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (r[j] >= 0 && r[j] < tgt) tgt = r[j];
  bool need[9], ok[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    need[j] = r[j] >= 0 && r[j] != tgt;
#pragma unroll
    for (int k = 0; k < j; ++k)
      if (r[k] == r[j]) need[j] = false;
  }
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    ok[j] = true;
    if (need[j]) {
      int expected = r[j];  // a root, larger than tgt
      ok[j] = __hip_atomic_compare_exchange_strong(parent + r[j], &expected, tgt, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  // units of the hooked roots of this level's elements, added to tgt; the
  // lanes of a wave adding to the first adding lane's target combine into one
  // atomic (sums by bit-sliced ballots, valid with inactive lanes)
  int cnt = 0;
#pragma unroll
  for (int j = 0; j < 9; ++j) cnt += (need[j] && ok[j] && r[j] >= b0) ? 1 : 0;
  const int tgt0 = tgt;
#pragma unroll
  for (int j = 0; j < 9; ++j)
    if (!ok[j]) tgt = uf_link(parent, size, r[j], tgt, b0);
  const unsigned long long am = __ballot(cnt > 0);
  if (am != 0ull) {
    const int leader = __ffsll((long long)am) - 1;
    const int tl = __shfl(tgt0, leader, 64);
    const bool same = cnt > 0 && tgt0 == tl;
    int tot = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) tot += __popcll(__ballot(same && ((cnt >> b) & 1))) << b;
    if ((threadIdx.x & 63) == leader)
      __hip_atomic_fetch_add(size + tl, tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else if (cnt > 0 && !same)
      __hip_atomic_fetch_add(size + tgt0, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  return tgt;
}

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// First level l with x > T[l] (T strictly decreasing, L levels), kNever if none.
__device__ __forceinline__ int level_of(float xv, const float* s_T, int L, float T0, float dT) {
  if (!(xv > s_T[L - 1])) return kNever;  // also NaN
  int l = 0;
  if (xv < T0) {
    const float g = (T0 - xv) / dT;  // estimate; fixed up against the table below
    l = g < (float)(L - 1) ? (int)g : L - 1;
    if (l < 0) l = 0;
  }
  while (l > 0 && xv > s_T[l - 1]) --l;
  while (!(xv > s_T[l])) ++l;
  return l;
}

struct ItemRef {
  int s, p;
  float sign;
};

__device__ __forceinline__ ItemRef decode_item(const TfceBatchArgs& a, int64_t item) {
  const int per_surf = a.n_maps * a.n_signs;
  ItemRef r;
  r.s = (int)(item / per_surf);
  const int rem = (int)(item - (int64_t)r.s * per_surf);
  r.p = rem / a.n_signs;
  r.sign = (a.n_signs == 2) ? ((rem % 2) ? -1.0f : 1.0f) : a.sign0;
  return r;
}

// fl(T^H), correctly rounded (T*T is exact in double).  The reference calls libm
// powf, which is one ulp off for ~0.06% of inputs.
__device__ __forceinline__ float pow_h(float T, float H) {
  if (H == 2.0f) return __fmul_rn(T, T);
  if (H == 1.0f) return T;
  return __double2float_rn(pow((double)T, (double)H));
}

// One workgroup per item: maximum of sign*map over the surface (NaN ignored),
// first threshold, step, and the level tables.
__global__ __launch_bounds__(kTfceThreads) void tfce_prepare_kernel(TfceBatchArgs a) {
  __shared__ float s_red[kTfceThreads / 64];
  const int64_t item = blockIdx.x;
  const ItemRef it = decode_item(a, item);
  const SurfDev S = a.surfs[it.s];
  const float* x = a.maps + (int64_t)it.p * a.map_stride + S.offset;
  const int V = S.V;
  float m0 = -INFINITY, m1 = -INFINITY, m2 = -INFINITY, m3 = -INFINITY;
  int v = threadIdx.x;
  for (; v + 3 * kTfceThreads < V; v += 4 * kTfceThreads) {
    m0 = fmaxf(m0, it.sign * x[v]);
    m1 = fmaxf(m1, it.sign * x[v + kTfceThreads]);
    m2 = fmaxf(m2, it.sign * x[v + 2 * kTfceThreads]);
    m3 = fmaxf(m3, it.sign * x[v + 3 * kTfceThreads]);
  }
  for (; v < V; v += kTfceThreads) m0 = fmaxf(m0, it.sign * x[v]);
  const float m = block_max(fmaxf(fmaxf(m0, m1), fmaxf(m2, m3)), s_red);
  if (threadIdx.x != 0) return;
  int32_t* rec = a.prep + item * kPrepWords;
  float* T = reinterpret_cast<float*>(rec + kPrepT);
  float* HH = reinterpret_cast<float*>(rec + kPrepT + kMaxLevels);
  const float T0 = a.ext_T0 ? a.ext_T0[item] : ((S.flags & kSurfZeroInLevels) ? fmaxf(m, 0.0f) : m);
  const float dT = (a.dh != 0.0f) ? a.dh : __fdiv_rn(T0, 100.0f);
  int L = 0, status = 0;
  if (T0 >= a.min_t) {
    if (!(dT > 0.0f)) {
      status = kStatusDegenerate;
    } else {
      float t = T0;
      while (t >= a.min_t) {
        if (L == kMaxLevels) { status = kStatusTooManyLevels; break; }
        T[L] = t;
        HH[L] = pow_h(t, S.H);
        ++L;
        t = __fsub_rn(t, dT);
      }
    }
  }
  rec[kPrepL] = status < 0 ? status : L;
  rec[kPrepOwn] = __float_as_int(m);
  rec[kPrepT0] = __float_as_int(T0);
  rec[kPrepDT] = __float_as_int(dT);
}

// RING: every surface of the launch is ring-packed (closed triangle meshes of
// at most six neighbours per vertex); only that path is compiled, which keeps
// the register allocation of the other table layouts out of the kernel.
template <bool FULL, bool RING>
__global__ __launch_bounds__(kTfceThreads, FULL ? 3 : 4) void tfce_batch_kernel(TfceBatchArgs a) {
  __shared__ float s_T[kMaxLevels];
  __shared__ float s_HH[kMaxLevels];
  __shared__ int s_start[kMaxLevels + 1];
  __shared__ int s_cur[kMaxLevels];
  __shared__ float s_red[kTfceThreads / 64];
  __shared__ int s_item_lo, s_item_hi, s_nroots_new, s_nact;

  const int tid = threadIdx.x;
  const int64_t vstride = (a.v_max + 63) & ~63;
  int32_t* slot = a.scratch + (int64_t)blockIdx.x * a.slot_stride;
  int32_t* pos = slot;
  int32_t* order = slot + vstride;
  int32_t* parent = slot + 2 * vstride;
  int32_t* size = slot + 3 * vstride;
  float* val = reinterpret_cast<float*>(slot + 4 * vstride);
  int32_t* roots[2] = {slot + 5 * vstride, slot + 6 * vstride};
  float* incv = reinterpret_cast<float*>(slot + 7 * vstride);
  int32_t* npos = slot + 8 * vstride;  // [element][W] neighbour positions (a.np_w > 0)

  while (true) {
    if (tid == 0) {
      unsigned long long it = atomicAdd(reinterpret_cast<unsigned long long*>(a.queue), 1ull);
      s_item_lo = (int)(it & 0x7fffffff);
      s_item_hi = (int)(it >> 31);
    }
    __syncthreads();
    const int64_t item = ((int64_t)s_item_hi << 31) | s_item_lo;
    if (item >= a.n_items) break;  // uniform across the workgroup

    const ItemRef itr = decode_item(a, item);
    const float sign = itr.sign;
    const SurfDev S = a.surfs[itr.s];
    const int V = S.V;
    const float* x = a.maps + (int64_t)itr.p * a.map_stride + S.offset;
    float* enhn = (FULL && a.enhn) ? a.enhn + (int64_t)itr.p * a.map_stride + S.offset : nullptr;
    const float* vd = (FULL && a.vdensity) ? a.vdensity + S.offset : nullptr;
    const bool symmetric = RING || (S.flags & kSurfSymmetric) != 0;
    const int32_t* ell = S.ell;
    const int W = S.ell_width;
    const int32_t* rec = a.prep + item * kPrepWords;
    const int status = rec[kPrepL];
    const int L = status > 0 ? status : 0;
    const float T0 = __int_as_float(rec[kPrepT0]);
    const float dT = __int_as_float(rec[kPrepDT]);
    const float own = __int_as_float(rec[kPrepOwn]);

    // ---- phase 1: level tables, activation level of every element ------------
    for (int l = tid; l < L; l += kTfceThreads) {
      s_T[l] = __int_as_float(rec[kPrepT + l]);
      s_HH[l] = __int_as_float(rec[kPrepT + kMaxLevels + l]);
      s_cur[l] = 0;
    }
    if (tid == 0) s_nact = 0;
    __syncthreads();
    // eight elements per thread and pass, loads before stores
    if (L > 0) {
      for (int v0 = tid; v0 < V; v0 += 8 * kTfceThreads) {
        float xs[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int v = v0 + j * kTfceThreads;
          xs[j] = v < V ? x[v] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int v = v0 + j * kTfceThreads;
          if (v < V) {
            const int lev = level_of(sign * xs[j], s_T, L, T0, dT);
            if (lev != kNever) atomicAdd(&s_cur[lev], 1);
          }
        }
      }
    }
    __syncthreads();
    // exclusive scan of the histogram by wave 0
    if (tid < 64) {
      const int per = (L + 63) / 64;
      const int b = tid * per, e = min(L, b + per);
      int sum = 0;
      for (int l = b; l < e; ++l) sum += s_cur[l];
      const int incl = wave_incl_scan(sum);
      int run = incl - sum;
      for (int l = b; l < e; ++l) {
        const int c = s_cur[l];
        s_start[l] = run;
        s_cur[l] = run;
        run += c;
      }
      if (tid == 63) s_start[L] = incl;
    }
    __syncthreads();
    // ---- phase 2: level order ----------------------------------------------
    // With neighbour-position rows the MAX sweep never reads the level order
    // (vertex of a position): only the positions are written then.
    const bool ring = RING || (S.flags & kSurfRing) != 0;
    const bool use_np = RING || (symmetric && ell != nullptr && (ring ? kRingRowWords : W) <= a.np_w);
    // wide symmetric tables: rows of joined positions (kCandRowWords)
    const bool use_cr = !RING && !use_np && symmetric && ell != nullptr && S.adjw != nullptr &&
                        W == kEllMaxWidth && a.np_w >= kCandRowWords;
    const bool need_order = FULL || !(use_np || use_cr);
    if (L > 0) {
      // the levels again from the map (cheaper than storing and re-reading
      // them); every element's position is written here, kNever if inactive
      for (int v0 = tid; v0 < V; v0 += 8 * kTfceThreads) {
        float xs[8];
        int lv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int v = v0 + j * kTfceThreads;
          xs[j] = v < V ? x[v] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int v = v0 + j * kTfceThreads;
          lv[j] = v < V ? level_of(sign * xs[j], s_T, L, T0, dT) : kNever;
          if (v < V && lv[j] == kNever) pos[v] = kNever;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if (lv[j] != kNever) {
            const int v = v0 + j * kTfceThreads;
            const int q = atomicAdd(&s_cur[lv[j]], 1);
            if (need_order) order[q] = v;
            pos[v] = q;
          }
        }
        // joined-position rows in MAX mode: the active elements, in element
        // order (appended per wave), so that the rows pass visits only them
        // (FULL mode keeps the level order there, which lists them too)
        if (use_cr && !need_order) {
#pragma unroll
          for (int j = 0; j < 8; ++j) wave_append(lv[j] != kNever, v0 + j * kTfceThreads, order, &s_nact);
        }
      }
    }
    if (tid == 0) s_nroots_new = 0;
    __syncthreads();
    const int A = (L > 0) ? s_start[L] : 0;
    // Symmetric surface with a fixed-width table: the positions of every
    // active element's neighbours, one row per element in position order.  The
    // pass runs in vertex order, where rows of the table and neighbours'
    // positions are near each other, and the union step then reads one row per
    // element, consecutive across a wave, instead of a table row and a gather.
    if (use_np && L > 0) {
      // four vertices per thread and pass, every load of a 4-wide column
      // chunk issued before its stores
      for (int v0 = tid; v0 < V; v0 += 4 * kTfceThreads) {
        int qv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int v = v0 + j * kTfceThreads;
          qv[j] = v < V ? pos[v] : kNever;
        }
        if (ring) {
          // one 16-byte row per element: the packed table row (neighbour
          // indices, links, length) in, the same row of positions out
          const int4* tab = reinterpret_cast<const int4*>(S.nadj);
          int4 tw[4];
#pragma unroll
          for (int t = 0; t < 4; ++t) tw[t] = qv[t] != kNever ? tab[v0 + t * kTfceThreads] : make_int4(-1, -1, -1, -1);
          int q[4][kRingMaxDeg];
          unsigned tail[4];
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            int u[kRingMaxDeg];
            ring_fields(tw[t], u, tail[t]);
#pragma unroll
            for (int j = 0; j < kRingMaxDeg; ++j) q[t][j] = u[j] != kRingNone ? pos[u[j]] : kNever;
          }
#pragma unroll
          for (int t = 0; t < 4; ++t)
            if (qv[t] != kNever) reinterpret_cast<int4*>(npos)[qv[t]] = ring_pack(q[t], tail[t]);
          continue;
        }
        if (W == 8) {
          // the common mesh width: both column chunks at once
          int4 uu[8], o[8];
#pragma unroll
          for (int j = 0; j < 8; ++j)
            uu[j] = qv[j >> 1] != kNever
                        ? reinterpret_cast<const int4*>(ell + (size_t)(v0 + (j >> 1) * kTfceThreads) * 8)[j & 1]
                        : make_int4(-1, -1, -1, -1);
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            o[j].x = uu[j].x >= 0 ? pos[uu[j].x] : kNever;
            o[j].y = uu[j].y >= 0 ? pos[uu[j].y] : kNever;
            o[j].z = uu[j].z >= 0 ? pos[uu[j].z] : kNever;
            o[j].w = uu[j].w >= 0 ? pos[uu[j].w] : kNever;
          }
          if (S.nadj) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              if (qv[j >> 1] == kNever) continue;
              const uint8_t* nr = S.nadj + (size_t)(v0 + (j >> 1) * kTfceThreads) * 8 + (j & 1) * 4;
              o[j].x = np_encode(o[j].x, nr, 0);
              o[j].y = np_encode(o[j].y, nr, 1);
              o[j].z = np_encode(o[j].z, nr, 2);
              o[j].w = np_encode(o[j].w, nr, 3);
            }
          }
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if (qv[j >> 1] != kNever) reinterpret_cast<int4*>(npos + (size_t)qv[j >> 1] * 8)[j & 1] = o[j];
          continue;
        }
        for (int c = 0; c < (W >> 2); ++c) {
          int4 uu[4], o[4];
#pragma unroll
          for (int j = 0; j < 4; ++j)
            uu[j] = qv[j] != kNever ? reinterpret_cast<const int4*>(ell + (size_t)(v0 + j * kTfceThreads) * W)[c]
                                    : make_int4(-1, -1, -1, -1);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            o[j].x = uu[j].x >= 0 ? pos[uu[j].x] : kNever;
            o[j].y = uu[j].y >= 0 ? pos[uu[j].y] : kNever;
            o[j].z = uu[j].z >= 0 ? pos[uu[j].z] : kNever;
            o[j].w = uu[j].w >= 0 ? pos[uu[j].w] : kNever;
          }
          if (S.nadj) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              if (qv[j] == kNever) continue;
              const uint8_t* nr = S.nadj + (size_t)(v0 + j * kTfceThreads) * W + c * 4;
              o[j].x = np_encode(o[j].x, nr, 0);
              o[j].y = np_encode(o[j].y, nr, 1);
              o[j].z = np_encode(o[j].z, nr, 2);
              o[j].w = np_encode(o[j].w, nr, 3);
            }
          }
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (qv[j] != kNever) reinterpret_cast<int4*>(npos + (size_t)qv[j] * W)[c] = o[j];
        }
      }
      __syncthreads();
    }
    // Wide symmetric table: the joined positions of every active element, in
    // element order over the list of active elements (table rows and
    // neighbours' positions near each other):
    // the neighbours activated earlier in its level, and one of each group of
    // mutually adjacent neighbours activated at earlier levels (one set
    // already), found with the row's full adjacency.  The union step then
    // reads one 16-byte row per element instead of a table row, its
    // adjacency and 26 positions.
    if (use_cr && L > 0) {
      constexpr int CW = kEllMaxWidth;
      for (int k = tid; k < A; k += kTfceThreads) {
        const int v = order[k];  // the k-th active element
        const int q = pos[v];
        // the table row (one line), then every neighbour's position and the
        // row's adjacency masks (one line): two rounds of independent loads
        const int4* trow4 = reinterpret_cast<const int4*>(ell + (size_t)v * CW);
        const size_t arow = S.adjid ? (size_t)S.adjid[v] : (size_t)v;
        const uint4* arow4 = reinterpret_cast<const uint4*>(S.adjw + arow * CW);
        int u[CW];
        unsigned m[CW];
#pragma unroll
        for (int c = 0; c < CW / 4; ++c) {
          const int4 t4 = trow4[c];
          u[4 * c] = t4.x; u[4 * c + 1] = t4.y; u[4 * c + 2] = t4.z; u[4 * c + 3] = t4.w;
        }
        int qu[CW];
#pragma unroll
        for (int j = 0; j < CW; ++j) qu[j] = u[j] >= 0 ? pos[u[j]] : kNever;
#pragma unroll
        for (int c = 0; c < CW / 4; ++c) {
          const uint4 a4 = arow4[c];
          m[4 * c] = a4.x; m[4 * c + 1] = a4.y; m[4 * c + 2] = a4.z; m[4 * c + 3] = a4.w;
        }
        // first position of v's level: the last level start <= q
        int lo = 0, hi = L;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_start[mid] <= q) lo = mid;
          else hi = mid;
        }
        const int b0v = s_start[lo];
        unsigned old_m = 0u, same_m = 0u;
#pragma unroll
        for (int j = 0; j < CW; ++j) {
          if (qu[j] < b0v) old_m |= 1u << j;
          else if (qu[j] < q) same_m |= 1u << j;
        }
        // one seed per connected group of the earlier levels' neighbours
        // (flood fill over the masks, one hop per round, in registers)
        unsigned reps = 0u, rem = old_m;
        while (rem) {
          const unsigned seed = rem & (0u - rem);
          unsigned comp = seed, prev;
          do {
            prev = comp;
            unsigned grow = 0u;
#pragma unroll
            for (int j = 0; j < CW; ++j) grow |= (comp >> j) & 1u ? m[j] : 0u;
            comp = (comp | grow) & old_m;
          } while (comp != prev);
          reps |= seed;
          rem &= ~comp;
        }
        const unsigned cand = reps | same_m;
        int4 w = make_int4(-1, -1, -1, -1);
        if (__popc(cand) > kCandRowWords) {
          w.x = -2 - v;  // joined from the table row
        } else {
          // the candidates' positions, in entry order, through a 4-slot shift
#pragma unroll
          for (int j = CW - 1; j >= 0; --j)
            if ((cand >> j) & 1u) {
              w.w = w.z;
              w.z = w.y;
              w.y = w.x;
              w.x = qu[j];
            }
        }
        reinterpret_cast<int4*>(npos)[q] = w;
      }
      __syncthreads();
    }

    // ---- phase 3: threshold sweep -------------------------------------------
    int n_roots = 0;
    int cur = 0;
    for (int l = 0; l < L; ++l) {
      const int b0 = s_start[l], b1 = s_start[l + 1];
      const float HH = s_HH[l];
      int32_t* rold = roots[cur];
      int32_t* rnew = roots[cur ^ 1];
      if (b1 > b0) {
        // 3a: one set per new vertex
        for (int i = b0 + tid; i < b1; i += kTfceThreads) {
          parent[i] = i;
          size[i] = 1;
          // (MAX mode: a value is first read as a root's at its own level's
          // increment, which takes 0 for roots of that level, so none is
          // written here)
          if (FULL) val[i] = enhn ? enhn[order[i]] : 0.0f;
        }
        __syncthreads();
        // 3b: join new vertices with their active neighbours, 8 at a time
        // (join_chunk).  With the fixed-width table a thread's next element's
        // first row and the vertex of the one after are loaded one iteration
        // ahead.
        const int4 none4 = make_int4(-1, -1, -1, -1);
        if (use_np && ring) {
          // one packed row per element, the next one loaded ahead
          const int4* rows = reinterpret_cast<const int4*>(npos);
          int4 w_n = (b0 + tid < b1) ? rows[b0 + tid] : none4;
          for (int i = b0 + tid; i < b1; i += kTfceThreads) {
            const int4 w = w_n;
            if (i + kTfceThreads < b1) w_n = rows[i + kTfceThreads];
            int p[8], r[9];
            unsigned adj[8];
            ring_unpack(w, p, adj);
            prune8(p, adj, i, b0, r);
            join_roots<false>(parent, size, r, i, b0);
          }
        } else if (use_cr) {
          // one row of joined positions per element, the next one loaded ahead
          const int4* rows = reinterpret_cast<const int4*>(npos);
          int4 w_n = (b0 + tid < b1) ? rows[b0 + tid] : none4;
          for (int i = b0 + tid; i < b1; i += kTfceThreads) {
            const int4 w = w_n;
            if (i + kTfceThreads < b1) w_n = rows[i + kTfceThreads];
            if (w.x <= -2) {
              join_table_row(parent, size, pos, ell + (size_t)(-2 - w.x) * W, W, i, b0);
              continue;
            }
            int r[9] = {w.x, w.y, w.z, w.w, -1, -1, -1, -1, 0};
            join_roots<false>(parent, size, r, i, b0);
          }
        } else if (use_np) {
          // rows of neighbour positions: a joined neighbour is one with a
          // smaller position (active before this level, or earlier in it)
          int4 lo_n = none4, hi_n = none4;
          if (b0 + tid < b1) {
            const int4* rn = reinterpret_cast<const int4*>(npos + (size_t)(b0 + tid) * W);
            lo_n = rn[0];
            if (W > 4) hi_n = rn[1];
          }
          for (int i = b0 + tid; i < b1; i += kTfceThreads) {
            const int4 lo = lo_n, hi = hi_n;
            if (i + kTfceThreads < b1) {
              const int4* rn = reinterpret_cast<const int4*>(npos + (size_t)(i + kTfceThreads) * W);
              lo_n = rn[0];
              if (W > 4) hi_n = rn[1];
            }
            int r[9] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w, 0};
            if (S.nadj) {
              const int e8[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
              prune_row(e8, i, b0, r);
            } else {
#pragma unroll
              for (int j = 0; j < 8; ++j) r[j] = r[j] < i ? r[j] : -1;
            }
            if (W <= 8) {
              join_roots<false>(parent, size, r, i, b0);
              continue;
            }
            // rows of 12 or 16: both chunks' candidates, sorted, split
            // bitonically into the 8 largest (all of them when they are 8 or
            // fewer) and the rest, joined only when a lane has any
            const int4* row = reinterpret_cast<const int4*>(npos + (size_t)i * W);
            const int4 l2 = row[2];
            const int4 h2 = (W > 12) ? row[3] : make_int4(kNever, kNever, kNever, kNever);
            int r2[9] = {l2.x, l2.y, l2.z, l2.w, h2.x, h2.y, h2.z, h2.w, 0};
            if (S.nadj) {
              const int e8[8] = {l2.x, l2.y, l2.z, l2.w, h2.x, h2.y, h2.z, h2.w};
              prune_row(e8, i, b0, r2);
            } else {
#pragma unroll
              for (int j = 0; j < 8; ++j) r2[j] = r2[j] < i ? r2[j] : -1;
            }
            sort8_desc(r);
            sort8_desc(r2);
            int top[9], bot[9];
            bool spill = false;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              top[j] = max(r[j], r2[7 - j]);
              bot[j] = min(r[j], r2[7 - j]);
              spill |= bot[j] >= 0;
            }
            const int ri = join_roots<false>(parent, size, top, i, b0);
            if (__any(spill)) join_roots<false>(parent, size, bot, ri, b0);
          }
        } else {
          int v_nxt = (b0 + tid < b1) ? order[b0 + tid] : 0;
          int v_nn = (b0 + tid + kTfceThreads < b1) ? order[b0 + tid + kTfceThreads] : 0;
          int4 lo_nxt = none4, hi_nxt = none4;
          if (ell && b0 + tid < b1) {
            const int4* rn = reinterpret_cast<const int4*>(ell + (size_t)v_nxt * W);
            lo_nxt = rn[0];
            if (W > 4) hi_nxt = rn[1];
          }
          for (int i = b0 + tid; i < b1; i += kTfceThreads) {
            const int v = v_nxt;
            const int4 lo0 = lo_nxt, hi0 = hi_nxt;
            if (ell && i + kTfceThreads < b1) {
              const int4* rn = reinterpret_cast<const int4*>(ell + (size_t)v_nn * W);
              lo_nxt = rn[0];
              if (W > 4) hi_nxt = rn[1];
            }
            v_nxt = v_nn;
            v_nn = (i + 2 * kTfceThreads < b1) ? order[i + 2 * kTfceThreads] : 0;
            const float xv = symmetric ? 0.0f : sign * x[v];
            int ri = i;  // root (or an ancestor) of i's set so far
            if (ell) {
              // no adjacency bytes (asymmetric lists): chunk by chunk
              int u[8] = {lo0.x, lo0.y, lo0.z, lo0.w, hi0.x, hi0.y, hi0.z, hi0.w};
              ri = join_chunk(parent, size, pos, x, u, i, v, ri, b0, b1, symmetric, sign, xv);
              // rows longer than 8: the table is -1 padded, so a chunk that ends
              // in padding is the last
              const int4* row = reinterpret_cast<const int4*>(ell + (size_t)v * W);
              for (int e = 8; e < W && u[7] >= 0; e += 8) {
                const int4 lo = row[e >> 2];
                const int4 hi = (e + 4 < W) ? row[(e >> 2) + 1] : none4;
                u[0] = lo.x; u[1] = lo.y; u[2] = lo.z; u[3] = lo.w;
                u[4] = hi.x; u[5] = hi.y; u[6] = hi.z; u[7] = hi.w;
                ri = join_chunk(parent, size, pos, x, u, i, v, ri, b0, b1, symmetric, sign, xv);
              }
            } else {
              const int e0 = S.row_ptr[v], e1 = S.row_ptr[v + 1];
              for (int e = e0; e < e1; e += 8) {
                int u[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) u[j] = (e + j < e1) ? S.col[e + j] : -1;
                ri = join_chunk(parent, size, pos, x, u, i, v, ri, b0, b1, symmetric, sign, xv);
              }
            }
          }
        }
        __syncthreads();
        // 3c: merge sizes (and MAX values) into surviving roots; collect roots.
        // An element of this level gave its own unit when it was hooked (3b);
        // what it received while it was a root (size - 1) goes to its final
        // root here, so the common case - hooked, never a root to anyone -
        // costs two consecutive loads and no find.  An old root that was
        // hooked passes its whole size and (MAX mode) its value.
        const int n_cand = n_roots + (b1 - b0);
        for (int k0 = 0; k0 < n_cand; k0 += kIlp * kTfceThreads) {
          int c[kIlp], r[kIlp], sz[kIlp];
          bool live[kIlp], fresh[kIlp], surv[kIlp], need[kIlp];
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            const int k = k0 + j * kTfceThreads + tid;
            live[j] = k < n_cand;
            fresh[j] = live[j] && k >= n_roots;
            c[j] = live[j] ? (fresh[j] ? b0 + (k - n_roots) : rold[k]) : 0;
          }
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            r[j] = fresh[j] ? ld(parent + c[j]) : c[j];
            sz[j] = fresh[j] ? ld(size + c[j]) : 0;
          }
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            surv[j] = fresh[j] && r[j] == c[j];  // old roots: known after the find
            need[j] = live[j] && (!fresh[j] || (r[j] != c[j] && sz[j] > 1));
          }
          find_batch(parent, r, need);
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            if (live[j] && !fresh[j]) {
              if (r[j] != c[j]) {
                __hip_atomic_fetch_add(size + r[j], ld(size + c[j]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (!FULL)
                  __hip_atomic_fetch_max(reinterpret_cast<int32_t*>(val) + r[j],
                                         ld(reinterpret_cast<const int32_t*>(val) + c[j]), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
              } else {
                surv[j] = true;
              }
            } else if (need[j]) {
              __hip_atomic_fetch_add(size + r[j], sz[j] - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            wave_append(surv[j], c[j], rnew, &s_nroots_new);
          }
        }
        __syncthreads();
        n_roots = s_nroots_new;
        __syncthreads();
        if (tid == 0) s_nroots_new = 0;
        cur ^= 1;
      }
      int32_t* rl = roots[cur];
      // 3d: level increment
      if (!FULL) {
        for (int k0 = 0; k0 < n_roots; k0 += kIlp * kTfceThreads) {
          int r[kIlp], sz[kIlp];
          float cv[kIlp];
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            const int k = k0 + j * kTfceThreads + tid;
            r[j] = k < n_roots ? rl[k] : -1;
          }
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            sz[j] = r[j] >= 0 ? ld(size + r[j]) : 0;
            cv[j] = (r[j] >= 0 && r[j] < b0) ? __int_as_float(ld(reinterpret_cast<const int32_t*>(val) + r[j])) : 0.0f;
          }
#pragma unroll
          for (int j = 0; j < kIlp; ++j)
            if (r[j] >= 0) val[r[j]] = cv[j] + tfce_inc(sz[j], HH, S.E, S.sizepow);
        }
        __syncthreads();
      } else {
        for (int k = tid; k < n_roots; k += kTfceThreads) {
          const int r = rl[k];
          incv[r] = tfce_inc(ld(size + r), HH, S.E, S.sizepow);
        }
        __syncthreads();
        for (int i0 = 0; i0 < b1; i0 += kIlp * kTfceThreads) {
          int r[kIlp];
          bool live[kIlp];
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            const int i = i0 + j * kTfceThreads + tid;
            live[j] = i < b1;
            r[j] = live[j] ? i : 0;
          }
          find_batch(parent, r, live);
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            const int i = i0 + j * kTfceThreads + tid;
            if (live[j]) {
              val[i] = val[i] + incv[r[j]];
              if (a.labels && l == a.labels_level) a.labels[order[i]] = order[r[j]];
            }
          }
        }
        __syncthreads();
      }
    }

    // ---- finalize -----------------------------------------------------------
    float best = 0.0f;
    if (!FULL) {
      int32_t* rl = roots[cur];
      for (int k = tid; k < n_roots; k += kTfceThreads)
        best = fmaxf(best, __int_as_float(ld(reinterpret_cast<const int32_t*>(val) + rl[k])));
    } else {
      for (int i = tid; i < A; i += kTfceThreads) best = fmaxf(best, val[i]);
      if (enhn)
        for (int i = tid; i < A; i += kTfceThreads) enhn[order[i]] = val[i];
    }
    best = block_max(best, s_red);
    const float smax = (S.flags & kSurfZeroInScale) ? fmaxf(own, 0.0f) : own;
    const float sc = __fdiv_rn(smax, 100.0f);
    float scaled;
    if (FULL && vd) {
      // max over all elements of (tfce * sc) * vd; inactive elements hold tfce = 0
      float mv = -INFINITY;
      for (int v = tid; v < V; v += kTfceThreads) {
        const int q = pos[v];
        const float t = (L > 0 && q != kNever) ? val[q] : 0.0f;
        mv = fmaxf(mv, __fmul_rn(__fmul_rn(t, sc), vd[v]));
      }
      scaled = block_max(mv, s_red);
    } else {
      scaled = __fmul_rn(best, sc);
    }
    if (tid == 0) {
      a.raw_max[item] = best;
      if (a.scaled_max) a.scaled_max[item * a.out_mul + a.out_add] = scaled;
      if (a.own_max) a.own_max[item] = own;
      if (a.status) a.status[item] = status;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Paired sweep (MAX mode, ring-packed surfaces, both signs of every map): one
// workgroup takes a map's +image and -image together.  The elements active
// in the two are disjoint (x > T >= 0 against -x > T >= 0), so they share one
// set of per-element arrays: positions of the + image's elements first (by
// level), then the - image's; a neighbour-position row keeps only neighbours
// of its own element's sign, so the union-find of the two images never
// meets.  One pass over the map assigns both images' levels and positions and
// one pass over the vertices writes both images' rows (the separate sweeps
// read the map and walk the vertices once per image); the levels of the two
// run side by side, each with its own thresholds.  Same results as
// tfce_batch_kernel<false, true>, bit for bit.
constexpr int kPairThreads = 512;

__global__ __launch_bounds__(kPairThreads, 2) void tfce_pair_kernel(TfceBatchArgs a) {
  __shared__ float s_T[2][kMaxLevels];
  __shared__ float s_HH[2][kMaxLevels];
  __shared__ int s_start[2][kMaxLevels + 1];
  __shared__ int s_cur[2][kMaxLevels];
  __shared__ float s_red[2][kPairThreads / 64];
  __shared__ int s_unit_lo, s_unit_hi, s_nroots_new;

  const int tid = threadIdx.x;
  const int64_t vstride = (a.v_max + 63) & ~63;
  int32_t* slot = a.scratch + (int64_t)blockIdx.x * a.slot_stride;
  int32_t* pos = slot;
  int32_t* parent = slot + 2 * vstride;
  int32_t* size = slot + 3 * vstride;
  float* val = reinterpret_cast<float*>(slot + 4 * vstride);
  int32_t* roots[2] = {slot + 5 * vstride, slot + 6 * vstride};
  int32_t* npos = slot + 8 * vstride;  // [position][4] ring-packed neighbour positions
  const int64_t n_units = a.n_items / 2;

  while (true) {
    if (tid == 0) {
      unsigned long long it = atomicAdd(reinterpret_cast<unsigned long long*>(a.queue), 1ull);
      s_unit_lo = (int)(it & 0x7fffffff);
      s_unit_hi = (int)(it >> 31);
    }
    __syncthreads();
    const int64_t unit = ((int64_t)s_unit_hi << 31) | s_unit_lo;
    if (unit >= n_units) break;  // uniform across the workgroup

    const int64_t item0 = 2 * unit;  // +image; item0 + 1 is the -image
    const ItemRef itr = decode_item(a, item0);
    const SurfDev S = a.surfs[itr.s];
    const int V = S.V;
    const float* x = a.maps + (int64_t)itr.p * a.map_stride + S.offset;
    int L[2], status[2];
    float T0[2], dT[2], own[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int32_t* rec = a.prep + (item0 + g) * kPrepWords;
      status[g] = rec[kPrepL];
      L[g] = status[g] > 0 ? status[g] : 0;
      T0[g] = __int_as_float(rec[kPrepT0]);
      dT[g] = __int_as_float(rec[kPrepDT]);
      own[g] = __int_as_float(rec[kPrepOwn]);
      for (int l = tid; l < L[g]; l += kPairThreads) {
        s_T[g][l] = __int_as_float(rec[kPrepT + l]);
        s_HH[g][l] = __int_as_float(rec[kPrepT + kMaxLevels + l]);
        s_cur[g][l] = 0;
      }
    }
    __syncthreads();
    // ---- levels of both images from one read of the map -----------------------
    auto level2 = [&](float xv, int& g) -> int {
      int lv = kNever;
      g = 0;
      if (L[0] > 0) lv = level_of(xv, s_T[0], L[0], T0[0], dT[0]);
      if (lv == kNever && L[1] > 0) {
        lv = level_of(-xv, s_T[1], L[1], T0[1], dT[1]);
        g = 1;
      }
      return lv;
    };
    for (int v0 = tid; v0 < V; v0 += 8 * kPairThreads) {
      float xs[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int v = v0 + j * kPairThreads;
        xs[j] = v < V ? x[v] : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int v = v0 + j * kPairThreads;
        if (v < V) {
          int g;
          const int lev = level2(xs[j], g);
          if (lev != kNever) atomicAdd(&s_cur[g][lev], 1);
        }
      }
    }
    __syncthreads();
    // exclusive scans of the two histograms (waves 0 and 1), the - image's
    // positions after all of the + image's
    if (tid < 128) {
      const int g = tid >> 6, lane = tid & 63;
      const int Lg = L[g];
      const int per = (Lg + 63) / 64;
      const int b = lane * per, e = min(Lg, b + per);
      int sum = 0;
      for (int l = b; l < e; ++l) sum += s_cur[g][l];
      const int incl = wave_incl_scan(sum);
      int run = incl - sum;
      for (int l = b; l < e; ++l) {
        const int c = s_cur[g][l];
        s_start[g][l] = run;
        s_cur[g][l] = run;
        run += c;
      }
      if (lane == 63) s_start[g][Lg] = incl;
    }
    __syncthreads();
    const int A0 = s_start[0][L[0]];  // the + image's active elements
    if (tid < 64) {
      for (int l = tid; l <= L[1]; l += 64) {
        s_start[1][l] += A0;
        if (l < L[1]) s_cur[1][l] += A0;
      }
    }
    __syncthreads();
    for (int v0 = tid; v0 < V; v0 += 8 * kPairThreads) {
      float xs[8];
      int lv[8], gs[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int v = v0 + j * kPairThreads;
        xs[j] = v < V ? x[v] : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int v = v0 + j * kPairThreads;
        lv[j] = v < V ? level2(xs[j], gs[j]) : kNever;
        if (v < V && lv[j] == kNever) pos[v] = kNever;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (lv[j] != kNever) pos[v0 + j * kPairThreads] = atomicAdd(&s_cur[gs[j]][lv[j]], 1);
    }
    if (tid == 0) s_nroots_new = 0;
    __syncthreads();
    // ---- one row of same-sign neighbour positions per active element ----------
    {
      const int4* tab = reinterpret_cast<const int4*>(S.nadj);
      for (int v0 = tid; v0 < V; v0 += 4 * kPairThreads) {
        int qv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int v = v0 + j * kPairThreads;
          qv[j] = v < V ? pos[v] : kNever;
        }
        int4 tw[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) tw[t] = qv[t] != kNever ? tab[v0 + t * kPairThreads] : make_int4(-1, -1, -1, -1);
        int q[4][kRingMaxDeg];
        unsigned tail[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          int u[kRingMaxDeg];
          ring_fields(tw[t], u, tail[t]);
          const bool neg = qv[t] >= A0;
#pragma unroll
          for (int j = 0; j < kRingMaxDeg; ++j) {
            const int qu = u[j] != kRingNone ? pos[u[j]] : kNever;
            q[t][j] = (qu != kNever && (qu >= A0) == neg) ? qu : kNever;
          }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
          if (qv[t] != kNever) reinterpret_cast<int4*>(npos)[qv[t]] = ring_pack(q[t], tail[t]);
      }
    }
    __syncthreads();
    // ---- the two threshold sweeps side by side ---------------------------------
    const int Lmax = max(L[0], L[1]);
    int n_roots = 0;
    int cur = 0;
    const int4* rows = reinterpret_cast<const int4*>(npos);
    const int4 none4 = make_int4(-1, -1, -1, -1);
    for (int l = 0; l < Lmax; ++l) {
      // this level's elements of each image: [b0[g], b1[g])
      int b0[2], b1[2];
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        b0[g] = l < L[g] ? s_start[g][l] : 0;
        b1[g] = l < L[g] ? s_start[g][l + 1] : 0;
      }
      const int n0 = b1[0] - b0[0], n_new = n0 + (b1[1] - b0[1]);
      int32_t* rold = roots[cur];
      int32_t* rnew = roots[cur ^ 1];
      if (n_new > 0) {
        for (int k = tid; k < n_new; k += kPairThreads) {
          const int i = k < n0 ? b0[0] + k : b0[1] + (k - n0);
          parent[i] = i;
          size[i] = 1;
        }
        __syncthreads();
        // join this level's elements with their joined neighbours
        auto elem = [&](int k, int& bb) -> int {
          bb = k < n0 ? b0[0] : b0[1];
          return k < n0 ? b0[0] + k : b0[1] + (k - n0);
        };
        int bb_n;
        int4 w_n = tid < n_new ? rows[elem(tid, bb_n)] : none4;
        for (int k = tid; k < n_new; k += kPairThreads) {
          int bb;
          const int i = elem(k, bb);
          const int4 w = w_n;
          if (k + kPairThreads < n_new) w_n = rows[elem(k + kPairThreads, bb_n)];
          int p[8], r[9];
          unsigned adj[8];
          ring_unpack(w, p, adj);
          prune8(p, adj, i, bb, r);
          join_roots<false>(parent, size, r, i, bb);
        }
        __syncthreads();
        // merge sizes and values into surviving roots; collect roots (as the
        // single-image sweep's merge step, over both images' candidates)
        const int n_cand = n_roots + n_new;
        for (int k0 = 0; k0 < n_cand; k0 += kIlp * kPairThreads) {
          int c[kIlp], r[kIlp], sz[kIlp];
          bool live[kIlp], fresh[kIlp], surv[kIlp], need[kIlp];
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            const int k = k0 + j * kPairThreads + tid;
            live[j] = k < n_cand;
            fresh[j] = live[j] && k >= n_roots;
            int bb;
            c[j] = live[j] ? (fresh[j] ? elem(k - n_roots, bb) : rold[k]) : 0;
          }
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            r[j] = fresh[j] ? ld(parent + c[j]) : c[j];
            sz[j] = fresh[j] ? ld(size + c[j]) : 0;
          }
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            surv[j] = fresh[j] && r[j] == c[j];
            need[j] = live[j] && (!fresh[j] || (r[j] != c[j] && sz[j] > 1));
          }
          find_batch(parent, r, need);
#pragma unroll
          for (int j = 0; j < kIlp; ++j) {
            if (live[j] && !fresh[j]) {
              if (r[j] != c[j]) {
                __hip_atomic_fetch_add(size + r[j], ld(size + c[j]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_max(reinterpret_cast<int32_t*>(val) + r[j],
                                       ld(reinterpret_cast<const int32_t*>(val) + c[j]), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
              } else {
                surv[j] = true;
              }
            } else if (need[j]) {
              __hip_atomic_fetch_add(size + r[j], sz[j] - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            wave_append(surv[j], c[j], rnew, &s_nroots_new);
          }
        }
        __syncthreads();
        n_roots = s_nroots_new;
        __syncthreads();
        if (tid == 0) s_nroots_new = 0;
        cur ^= 1;
      }
      int32_t* rl = roots[cur];
      // level increment of every root of an image that still has levels; a
      // root of this level's elements starts from 0
      const float HH0 = l < L[0] ? s_HH[0][l] : 0.0f, HH1 = l < L[1] ? s_HH[1][l] : 0.0f;
      for (int k0 = 0; k0 < n_roots; k0 += kIlp * kPairThreads) {
        int r[kIlp], sz[kIlp];
        float cv[kIlp];
        bool go[kIlp];
#pragma unroll
        for (int j = 0; j < kIlp; ++j) {
          const int k = k0 + j * kPairThreads + tid;
          r[j] = k < n_roots ? rl[k] : -1;
          const int g = r[j] >= A0 ? 1 : 0;
          go[j] = r[j] >= 0 && l < L[g];
        }
#pragma unroll
        for (int j = 0; j < kIlp; ++j) {
          const int g = r[j] >= A0 ? 1 : 0;
          sz[j] = go[j] ? ld(size + r[j]) : 0;
          cv[j] = (go[j] && r[j] < b0[g]) ? __int_as_float(ld(reinterpret_cast<const int32_t*>(val) + r[j])) : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < kIlp; ++j)
          if (go[j]) val[r[j]] = cv[j] + tfce_inc(sz[j], r[j] >= A0 ? HH1 : HH0, S.E, S.sizepow);
      }
      __syncthreads();
    }

    // ---- finalize both items ---------------------------------------------------
    float best[2] = {0.0f, 0.0f};
    {
      int32_t* rl = roots[cur];
      for (int k = tid; k < n_roots; k += kPairThreads) {
        const int r = rl[k];
        const float vr = __int_as_float(ld(reinterpret_cast<const int32_t*>(val) + r));
        if (r < A0) best[0] = fmaxf(best[0], vr);
        else best[1] = fmaxf(best[1], vr);
      }
    }
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      float v = wave_max(best[g]);
      if ((tid & 63) == 0) s_red[g][tid >> 6] = v;
    }
    __syncthreads();
    if (tid < 2) {
      const int g = tid;
      float b = s_red[g][0];
      for (int w = 1; w < kPairThreads / 64; ++w) b = fmaxf(b, s_red[g][w]);
      const int64_t item = item0 + g;
      const float smax = (S.flags & kSurfZeroInScale) ? fmaxf(own[g], 0.0f) : own[g];
      const float sc = __fdiv_rn(smax, 100.0f);
      a.raw_max[item] = b;
      if (a.scaled_max) a.scaled_max[item * a.out_mul + a.out_add] = __fmul_rn(b, sc);
      if (a.own_max) a.own_max[item] = own[g];
      if (a.status) a.status[item] = status[g];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// LDS-resident sweep (MAX mode) for symmetric surfaces of at most kLdsMaxV
// elements.  One 1,024-thread workgroup per compute unit takes one map of a
// surface with both of its signs: the elements active in +map and in -map
// are disjoint (x > T >= 0 against -x > T >= 0), so the two images share one
// set of per-element arrays in LDS - activation position (the + image's
// elements first, then the - image's) and level order (16 bits each), one
// 32-bit union-find word per element (parent in the low 16 bits, set size in
// the high 16) and the value per element - and the two threshold sweeps run
// level by level through the same barriers.  Finds, hooks and size merges are
// LDS operations; only the map and the neighbour rows (shared by every image
// of the surface, cache resident) come from global memory.  Per level:
//   union     new elements join their active neighbours of the same image
//             (hook the larger root under the smaller by CAS on the parent
//             bits; sizes do not change in this phase, so a CAS fails only
//             when its root was hooked meanwhile);
//   merge     every element hooked at this level passes the size it held
//             (1 for a new element, the whole set for an old root) and its
//             value (atomicMax) to its final root, then points straight at it
//             with size 0 (so it is never counted twice); meanwhile each
//             thread gathers and prunes the neighbour candidates of its element
//             of the next level (they depend only on the positions);
//   increment every root adds (float)(size^E * T^H); the next level's
//             elements are initialised in the same pass.
// Same arithmetic as tfce_batch_kernel<false>: identical maxima, bit for bit.
constexpr int kLdsThreads = 1024;

struct LdsSweepShared {
  float T[2][kLdsMaxLevels];
  float HH[2][kLdsMaxLevels];
  int start[2][kLdsMaxLevels + 2];  // level l of image g: positions [base[g] + start[g][l], base[g] + start[g][l+1])
  int cur[2][kLdsMaxLevels];
  int base[2];
  float red[2][kLdsThreads / 64];
  int unit_lo, unit_hi;
  uint16_t pos[kLdsMaxV];    // element -> position in level order, 0xffff: never active
  uint16_t order[kLdsMaxV];  // position -> element
  uint32_t word[kLdsMaxV];   // parent | size << 16
  float val[kLdsMaxV];       // MAX value of the set a root heads
};

__device__ __forceinline__ uint32_t lds_ld(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_st(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Root of x.  SPLIT re-points a visited node at its grandparent, keeping its
// size bits: only safe while no size changes (the union phase).
template <bool SPLIT>
__device__ __forceinline__ int lds_find(uint32_t* w, int x) {
  uint32_t wx = lds_ld(w + x);
  int p = (int)(wx & 0xffffu);
  while (p != x) {
    const uint32_t wp = lds_ld(w + p);
    const int gp = (int)(wp & 0xffffu);
    if (SPLIT && gp != p) lds_st(w + x, (wx & 0xffff0000u) | (uint32_t)gp);
    x = p;
    wx = wp;
    p = gp;
  }
  return x;
}

// Join the sets of a and b (element positions or ancestors); returns the
// joined root (an ancestor of both afterwards).
__device__ __forceinline__ int lds_link(uint32_t* w, int a, int b) {
  while (true) {
    a = lds_find<true>(w, a);
    b = lds_find<true>(w, b);
    if (a == b) return a;
    const int hi = max(a, b), lo = min(a, b);
    uint32_t old = lds_ld(w + hi);
    if ((int)(old & 0xffffu) != hi) continue;  // hooked meanwhile
    if (__hip_atomic_compare_exchange_strong(w + hi, &old, (old & 0xffff0000u) | (uint32_t)lo, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
      return lo;
  }
}

// Position of element u in the image whose positions start at lo, kNever when
// u is not active in that image.
__device__ __forceinline__ int lds_pos(const uint16_t* pos, int u, int lo) {
  const int q = pos[u];
  return (q == 0xffff || q < lo) ? kNever : q;
}

// Candidates of element position i (image positions from lo, level from b0)
// of a ring-packed surface, from its packed row: one neighbour per group of
// mutually adjacent old neighbours, every joined neighbour of this level
// (prune8).
__device__ __forceinline__ void lds_ring_cands_row(const LdsSweepShared& sh, const int4 row, int i, int lo, int b0,
                                                   int (&r)[9]) {
  int u[kRingMaxDeg];
  unsigned tail;
  ring_fields(row, u, tail);
  int q[8];
#pragma unroll
  for (int j = 0; j < kRingMaxDeg; ++j) q[j] = u[j] != kRingNone ? lds_pos(sh.pos, u[j], lo) : kNever;
  q[6] = q[7] = kNever;
  unsigned adj[8];
  const unsigned link = tail & 63u;
  const int deg = (int)(tail >> 6);
#pragma unroll
  for (int j = 0; j < 8; ++j) adj[j] = 0;
#pragma unroll
  for (int j = 0; j < 5; ++j)
    if ((link >> j) & 1u) {
      adj[j] |= 2u << j;
      adj[j + 1] |= 1u << j;
    }
  if (link & 32u) {
#pragma unroll
    for (int j = 2; j < kRingMaxDeg; ++j)
      if (j == deg - 1) {
        adj[j] |= 1u;
        adj[0] |= 1u << j;
      }
  }
  prune8(q, adj, i, b0, r);
}

__device__ __forceinline__ void lds_ring_cands(const LdsSweepShared& sh, const SurfDev& S, int i, int lo, int b0,
                                               int (&r)[9]) {
  lds_ring_cands_row(sh, reinterpret_cast<const int4*>(S.nadj)[sh.order[i]], i, lo, b0, r);
}

// Union step of element position i (image positions from lo, level from b0)
// of a surface with a fixed-width table (pruned per group of 8 when adjacency
// bytes exist) or with the CSR only.
__device__ __forceinline__ void lds_join_table(LdsSweepShared& sh, const SurfDev& S, int i, int lo, int b0) {
  const int v = sh.order[i];
  int ri = i;
  if (S.ell) {
    const int W = S.ell_width;
    const int4* row = reinterpret_cast<const int4*>(S.ell + (size_t)v * W);
    for (int e = 0; e < W; e += 8) {
      const int4 l4 = row[e >> 2];
      const int4 h4 = (e + 4 < W) ? row[(e >> 2) + 1] : make_int4(-1, -1, -1, -1);
      const int uu[8] = {l4.x, l4.y, l4.z, l4.w, h4.x, h4.y, h4.z, h4.w};
      int q[8], r[9];
#pragma unroll
      for (int j = 0; j < 8; ++j) q[j] = uu[j] >= 0 ? lds_pos(sh.pos, uu[j], lo) : kNever;
      if (S.nadj) {
        unsigned wlo = 0, whi = 0, adj[8];
        chunk_adj(S.nadj, v, W, e, wlo, whi);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          adj[j] = (wlo >> (8 * j)) & 0xffu;
          adj[4 + j] = (whi >> (8 * j)) & 0xffu;
        }
        prune8(q, adj, i, b0, r);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) r[j] = q[j] < i ? q[j] : -1;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (r[j] >= 0) ri = lds_link(sh.word, ri, r[j]);
      if (uu[7] < 0) break;  // -1 padding: the row has ended
    }
  } else {
    for (int e = S.row_ptr[v]; e < S.row_ptr[v + 1]; ++e) {
      const int q = lds_pos(sh.pos, S.col[e], lo);
      if (q < i) ri = lds_link(sh.word, ri, q);
    }
  }
}

// The k-th new element of level l over both images: its position, the first
// position of its image and of its level.  n0 = new elements of image 0.
struct LdsNew {
  int i, lo, b0;
};
__device__ __forceinline__ LdsNew lds_new(const LdsSweepShared& sh, int k, int n0, int l) {
  LdsNew e;
  const int g = k < n0 ? 0 : 1;
  e.lo = sh.base[g];
  e.b0 = e.lo + sh.start[g][l];
  e.i = e.b0 + (g ? k - n0 : k);
  return e;
}

__global__ __launch_bounds__(kLdsThreads, 1) void tfce_lds_kernel(TfceBatchArgs a, int64_t unit0) {
  __shared__ LdsSweepShared sh;
  const int tid = threadIdx.x;
  const int ng = a.n_signs;
  unsigned long long* queue = reinterpret_cast<unsigned long long*>(a.queue) + 1;
  const int64_t n_units = a.n_items / ng;
  while (true) {
    if (tid == 0) {
      const unsigned long long u = atomicAdd(queue, 1ull) + (unsigned long long)unit0;
      sh.unit_lo = (int)(u & 0x7fffffff);
      sh.unit_hi = (int)(u >> 31);
    }
    __syncthreads();
    const int64_t unit = ((int64_t)sh.unit_hi << 31) | sh.unit_lo;
    if (unit >= n_units) break;  // uniform across the workgroup
    const int64_t item0 = unit * ng;
    const ItemRef itr = decode_item(a, item0);
    const SurfDev S = a.surfs[itr.s];
    const int V = S.V;
    const float* x = a.maps + (int64_t)itr.p * a.map_stride + S.offset;
    const bool ring = (S.flags & kSurfRing) != 0;
    int Lg[2] = {0, 0};
    float T0g[2] = {0.0f, 0.0f}, dTg[2] = {1.0f, 1.0f}, sgn[2] = {1.0f, -1.0f};
    bool bad = false;
    for (int g = 0; g < ng; ++g) {
      const int32_t* rec = a.prep + (item0 + g) * kPrepWords;
      const int st = rec[kPrepL];
      Lg[g] = st > 0 ? st : 0;
      if (Lg[g] > kLdsMaxLevels) {
        bad = true;
        Lg[g] = 0;
      }
      T0g[g] = __int_as_float(rec[kPrepT0]);
      dTg[g] = __int_as_float(rec[kPrepDT]);
      sgn[g] = decode_item(a, item0 + g).sign;
      for (int l = tid; l < Lg[g]; l += kLdsThreads) {
        sh.T[g][l] = __int_as_float(rec[kPrepT + l]);
        sh.HH[g][l] = __int_as_float(rec[kPrepT + kMaxLevels + l]);
        sh.cur[g][l] = 0;
      }
    }
    const int Lmax = max(Lg[0], Lg[1]);
    __syncthreads();
    // activation levels: one histogram per image
    for (int v = tid; v < V; v += kLdsThreads) {
      const float xv = x[v];
      for (int g = 0; g < ng; ++g) {
        if (Lg[g] == 0) continue;
        const int lev = level_of(sgn[g] * xv, sh.T[g], Lg[g], T0g[g], dTg[g]);
        if (lev != kNever) {
          atomicAdd(&sh.cur[g][lev], 1);
          break;
        }
      }
    }
    __syncthreads();
    if (tid < 64 * ng) {
      const int g = tid >> 6, lane = tid & 63;
      const int L = Lg[g];
      const int per = (L + 63) / 64;
      const int b = lane * per, e = min(L, b + per);
      int sum = 0;
      for (int l = b; l < e; ++l) sum += sh.cur[g][l];
      const int incl = wave_incl_scan(sum);
      int run = incl - sum;
      for (int l = b; l < e; ++l) {
        const int c = sh.cur[g][l];
        sh.start[g][l] = run;
        sh.cur[g][l] = run;
        run += c;
      }
      if (lane == 63) {
        sh.start[g][L] = incl;
        sh.start[g][L + 1] = incl;
      }
    }
    if (tid == 0) {
      sh.base[0] = 0;
      if (ng == 1) {
        sh.start[1][0] = 0;
        sh.start[1][1] = 0;
      }
    }
    __syncthreads();
    if (tid == 0) sh.base[1] = sh.start[0][Lg[0]];
    __syncthreads();
    for (int v = tid; v < V; v += kLdsThreads) {
      const float xv = x[v];
      int q = 0xffff;
      for (int g = 0; g < ng; ++g) {
        if (Lg[g] == 0) continue;
        const int lev = level_of(sgn[g] * xv, sh.T[g], Lg[g], T0g[g], dTg[g]);
        if (lev != kNever) {
          q = sh.base[g] + atomicAdd(&sh.cur[g][lev], 1);
          sh.order[q] = (uint16_t)v;
          break;
        }
      }
      sh.pos[v] = (uint16_t)q;
    }
    // the first level's elements (level 0 holds none: nothing exceeds the maximum)
    for (int g = 0; g < ng; ++g)
      for (int i = sh.base[g] + tid; i < sh.base[g] + (Lg[g] > 0 ? sh.start[g][1] : 0); i += kLdsThreads) {
        sh.word[i] = (1u << 16) | (uint32_t)i;
        sh.val[i] = 0.0f;
      }
    __syncthreads();

    // candidates of this thread's first new element of the next level, gathered
    // during the merge phase (they depend only on the positions; ring surfaces)
    int pre[9];
    bool have_pre = false;
    auto gather_next = [&](int l) {
      have_pre = false;
      if (!ring || l >= Lmax) return;
      const int n0 = l < Lg[0] ? sh.start[0][l + 1] - sh.start[0][l] : 0;
      const int n1 = l < Lg[1] ? sh.start[1][l + 1] - sh.start[1][l] : 0;
      if (tid < n0 + n1) {
        const LdsNew e = lds_new(sh, tid, n0, l);
        lds_ring_cands(sh, S, e.i, e.lo, e.b0, pre);
        have_pre = true;
      }
    };
    gather_next(0);
    for (int l = 0; l < Lmax; ++l) {
      const int n0 = l < Lg[0] ? sh.start[0][l + 1] - sh.start[0][l] : 0;
      const int n1 = l < Lg[1] ? sh.start[1][l + 1] - sh.start[1][l] : 0;
      // union
      for (int k = tid; k < n0 + n1; k += kLdsThreads) {
        const LdsNew e = lds_new(sh, k, n0, l);
        if (ring) {
          int r[9];
          if (k == tid && have_pre) {
#pragma unroll
            for (int j = 0; j < 9; ++j) r[j] = pre[j];
          } else {
            lds_ring_cands(sh, S, e.i, e.lo, e.b0, r);
          }
          int ri = e.i;
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if (r[j] >= 0) ri = lds_link(sh.word, ri, r[j]);
        } else {
          lds_join_table(sh, S, e.i, e.lo, e.b0);
        }
      }
      __syncthreads();
      // merge the sets hooked at this level into their roots (images still sweeping)
      const int m0 = l < Lg[0] ? sh.start[0][l + 1] : 0;
      const int m1 = l < Lg[1] ? sh.start[1][l + 1] : 0;
      const int b1base = sh.base[1];
      for (int k = tid; k < m0 + m1; k += kLdsThreads) {
        const int c = k < m0 ? k : b1base + (k - m0);
        const uint32_t wc = lds_ld(sh.word + c);
        if ((int)(wc & 0xffffu) == c || (wc >> 16) == 0u) continue;
        const int r = lds_find<false>(sh.word, (int)(wc & 0xffffu));
        __hip_atomic_fetch_add(sh.word + r, wc & 0xffff0000u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_max(reinterpret_cast<int32_t*>(sh.val) + r, __float_as_int(sh.val[c]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        lds_st(sh.word + c, (uint32_t)r);
      }
      gather_next(l + 1);
      __syncthreads();
      // increment every root of the images still sweeping; initialise the next level
      for (int k = tid; k < m0 + m1; k += kLdsThreads) {
        const int g = k < m0 ? 0 : 1;
        const int c = g ? b1base + (k - m0) : k;
        const uint32_t wc = sh.word[c];
        if ((int)(wc & 0xffffu) == c)
          sh.val[c] = sh.val[c] + tfce_inc((int)(wc >> 16), sh.HH[g][l], S.E, S.sizepow);
      }
      for (int g = 0; g < ng; ++g) {
        if (l + 1 >= Lg[g]) continue;
        const int lo = sh.base[g] + sh.start[g][l + 1], hi = sh.base[g] + sh.start[g][l + 2];
        for (int i = lo + tid; i < hi; i += kLdsThreads) {
          sh.word[i] = (1u << 16) | (uint32_t)i;
          sh.val[i] = 0.0f;
        }
      }
      __syncthreads();
    }

    // the largest value of each image (a root holds its set's largest)
    for (int g = 0; g < ng; ++g) {
      const int lo = sh.base[g], hi = lo + (Lg[g] > 0 ? sh.start[g][Lg[g]] : 0);
      float best = 0.0f;
      for (int c = lo + tid; c < hi; c += kLdsThreads) best = fmaxf(best, sh.val[c]);
      best = wave_max(best);
      if ((tid & 63) == 0) sh.red[g][tid >> 6] = best;
    }
    __syncthreads();
    if (tid < ng) {
      const int g = tid;
      float best = sh.red[g][0];
      for (int k = 1; k < kLdsThreads / 64; ++k) best = fmaxf(best, sh.red[g][k]);
      const int64_t item = item0 + g;
      const int32_t* rec = a.prep + item * kPrepWords;
      const float own = __int_as_float(rec[kPrepOwn]);
      const float smax = (S.flags & kSurfZeroInScale) ? fmaxf(own, 0.0f) : own;
      const float sc = __fdiv_rn(smax, 100.0f);
      const float raw = bad ? __int_as_float(0x7fc00000) : best;
      a.raw_max[item] = raw;
      if (a.scaled_max) a.scaled_max[item * a.out_mul + a.out_add] = __fmul_rn(raw, sc);
      if (a.own_max) a.own_max[item] = own;
      if (a.status) a.status[item] = bad ? kStatusTooManyLevels : rec[kPrepL];
    }
    __syncthreads();
  }
}

// Shared first threshold per (map, sign): the maximum over every surface of the
// set, as when the surfaces are TFCE'd as one merged image (tm_func.py:54-123,
// fast_tfce.hpp:34-35 then sees the global maximum).  One workgroup per
// (map, sign); writes the value to every surface's item.
__global__ __launch_bounds__(kTfceThreads) void shared_T0_kernel(TfceBatchArgs a, float* T0, int zero_in_image) {
  __shared__ float s_red[kTfceThreads / 64];
  const int p = blockIdx.x / a.n_signs, sg = blockIdx.x % a.n_signs;
  const float sign = (a.n_signs == 2) ? (sg ? -1.0f : 1.0f) : a.sign0;
  float m = -INFINITY;
  for (int s = 0; s < a.n_surf; ++s) {
    const SurfDev S = a.surfs[s];
    const float* x = a.maps + (int64_t)p * a.map_stride + S.offset;
    for (int v = threadIdx.x; v < S.V; v += kTfceThreads) m = fmaxf(m, sign * x[v]);
  }
  m = block_max(m, s_red);
  if (zero_in_image) m = fmaxf(m, 0.0f);
  for (int s = threadIdx.x; s < a.n_surf; s += kTfceThreads)
    T0[((int64_t)s * a.n_maps + p) * a.n_signs + sg] = m;
}

__global__ void sizepow_kernel(double* out, int32_t n, double E) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const double s = (double)i;
  out[i] = (E == 0.5) ? sqrt(s) : pow(s, E);
}

__global__ void fill_f32_kernel(float* p, int64_t n, float v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void zero_queue_kernel(int32_t* q) {
  q[0] = 0;
  q[1] = 0;
  q[2] = 0;
  q[3] = 0;
}

// ---------------------------------------------------------------------------
// One image over the whole GPU (FULL mode, symmetric adjacency): the drop-in
// CreateAdjSet.run on a large surface.  A cooperative grid (every workgroup
// resident) runs the levels in lock step, two grid-wide barriers per level:
//   X(l): every element of level l is joined with its active neighbours
//         (device-scope CAS union-find, larger root hooked under the
//         smaller); beside it, every element active at level l-1 adds that
//         level's increment of its component (float)(s^E * fl(T^H)), in
//         level order as fast_tfce.hpp:72-84 adds them (roots and sizes of
//         level l-1 were fixed in Y(l-1), and joins change neither);
//   Y(l): every new element adds one to its final root's size, every root
//         of the previous level that was hooked adds its whole size to its
//         final root, the surviving roots form the next root list, and every
//         active element's final root is recorded by position.
// TFCE values are accumulated by position (coalesced) and written back to the
// elements at the end; position i is always visited by the same thread.
__device__ __forceinline__ int gld(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gst(int32_t* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int guf_find(int32_t* parent, int x) {
  int p = gld(parent + x);
  while (p != x) {
    const int gp = gld(parent + p);
    if (gp != p) gst(parent + x, gp);  // path halving; gp is an ancestor of x
    x = p;
    p = gp;
  }
  return x;
}

// false if the join did not settle within the retry bound (the kernel then
// reports an error status instead of looping)
__device__ __forceinline__ bool guf_link(int32_t* parent, int a, int b) {
  for (int it = 0; it < (1 << 20); ++it) {
    a = guf_find(parent, a);
    b = guf_find(parent, b);
    if (a == b) return true;
    if (a < b) { const int t = a; a = b; b = t; }
    int expected = a;
    if (__hip_atomic_compare_exchange_strong(parent + a, &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return true;
  }
  return false;
}

// Grid-wide barrier of a cooperative launch (every workgroup is resident, so
// every wait ends): an arrival counter and a generation word at device scope;
// the last workgroup to arrive resets the counter and opens the next
// generation.  Workgroup-scope barriers and agent-scope fences around it make
// every workgroup's writes before it visible to every workgroup after it.  A
// wait that exceeds its bound (seconds) sets *fail and goes on.
__device__ __forceinline__ void grid_barrier(int32_t* bar, int32_t* fail, int& gen) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const int n = (int)gridDim.x;
    const int arrived = __hip_atomic_fetch_add(bar, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == n - 1) {
      __hip_atomic_store(bar, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(bar + 1, gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      int spins = 0;
      while (__hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1 << 25)) {
          __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  ++gen;
  __syncthreads();
}

constexpr int kCoopThreads = 256;

__global__ __launch_bounds__(kCoopThreads) void tfce_coop_kernel(TfceCoopArgs a) {
  __shared__ float s_T[kMaxLevels];
  __shared__ float s_HH[kMaxLevels];
  __shared__ int s_start[kMaxLevels + 1];
  const int tid = threadIdx.x;
  const int status = a.rec[kPrepL];
  if (blockIdx.x == 0 && tid == 0 && a.status && status <= 0) a.status[0] = status;
  const int L = status > 0 ? status : 0;
  if (L == 0) return;  // the same record for every workgroup: uniform
  const float T0 = __int_as_float(a.rec[kPrepT0]);
  const float dT = __int_as_float(a.rec[kPrepDT]);
  for (int l = tid; l < L; l += kCoopThreads) {
    s_T[l] = __int_as_float(a.rec[kPrepT + l]);
    s_HH[l] = __int_as_float(a.rec[kPrepT + kMaxLevels + l]);
  }
  __syncthreads();
  const int V = a.V;
  const int gtid = blockIdx.x * kCoopThreads + tid, gsz = gridDim.x * kCoopThreads;
  int32_t* start = a.ctl;                 // [kMaxLevels + 1]
  int32_t* cur = start + kMaxLevels + 1;  // [kMaxLevels]
  int32_t* nroots = cur + kMaxLevels;     // [3] root-list lengths, cycled by level
  int32_t* fail = nroots + 3;             // [1] a join or a wait that did not settle
  int32_t* bar = a.ctl + kCoopBarWord;    // [2] barrier counter and generation, zero at launch
  int gen = 0;
  for (int l = gtid; l < L; l += gsz) cur[l] = 0;
  if (gtid == 0) {
    nroots[0] = nroots[1] = nroots[2] = 0;
    *fail = 0;
  }
  grid_barrier(bar, fail, gen);
  // activation level of every element, histogram
  for (int v = gtid; v < V; v += gsz) {
    const int lv = level_of(a.x[v], s_T, L, T0, dT);
    a.lev[v] = lv;
    a.parent[v] = v;
    a.size[v] = 0;
    if (lv != kNever) atomicAdd(&cur[lv], 1);
  }
  grid_barrier(bar, fail, gen);
  if (blockIdx.x == 0 && tid < 64) {
    const int per = (L + 63) / 64;
    const int b = tid * per, e = min(L, b + per);
    int sum = 0;
    for (int l = b; l < e; ++l) sum += cur[l];
    const int incl = wave_incl_scan(sum);
    int run = incl - sum;
    for (int l = b; l < e; ++l) {
      const int c = cur[l];
      start[l] = run;
      cur[l] = run;
      run += c;
    }
    if (tid == 63) start[L] = incl;
  }
  grid_barrier(bar, fail, gen);
  for (int v = gtid; v < V; v += gsz) {
    const int lv = a.lev[v];
    if (lv != kNever) {
      const int q = atomicAdd(&cur[lv], 1);
      a.order[q] = v;
      a.acc[q] = a.enhn[v];
    }
  }
  for (int l = tid; l <= L; l += kCoopThreads) s_start[l] = start[l];
  grid_barrier(bar, fail, gen);
  const int32_t* rp = a.row_ptr;
  const int32_t* col = a.col;
  bool ok = true;
  for (int l = 0; l <= L; ++l) {
    // X(l): level l-1's increments, level l's joins
    if (l > 0) {
      // four positions per thread and pass, loads before the dependent ones
      const float HH = s_HH[l - 1];
      const int e0 = s_start[l];
      for (int i0 = gtid; i0 < e0; i0 += 4 * gsz) {
        int r[4], sz[4];
        float ac[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int i = i0 + j * gsz;
          r[j] = i < e0 ? a.rootc[i] : -1;
          ac[j] = i < e0 ? a.acc[i] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) sz[j] = r[j] >= 0 ? a.size[r[j]] : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int i = i0 + j * gsz;
          if (i < e0) a.acc[i] = ac[j] + tfce_inc(sz[j], HH, a.E, a.sizepow);
        }
      }
    }
    if (l == L) break;
    const int b = s_start[l], e = s_start[l + 1];
    for (int i = b + gtid; i < e; i += gsz) {
      const int v = a.order[i];
      for (int j = rp[v]; j < rp[v + 1]; ++j) {
        const int u = col[j];
        if (u != v && a.lev[u] <= l) ok = guf_link(a.parent, v, u) && ok;
      }
    }
    if (gtid == 0) nroots[(l + 2) % 3] = 0;
    grid_barrier(bar, fail, gen);
    // Y(l): sizes of the surviving roots, next root list, roots by position
    const int32_t* rc = a.roots[l & 1];
    int32_t* rn = a.roots[(l + 1) & 1];
    const int ncur = nroots[l % 3];
    for (int i = gtid; i < ncur; i += gsz) {
      const int r = rc[i];
      const int f = guf_find(a.parent, r);
      if (f == r) rn[atomicAdd(&nroots[(l + 1) % 3], 1)] = r;
      else atomicAdd(&a.size[f], a.size[r]);
    }
    // an older element's root changes only if that root was hooked
    for (int i0 = gtid; i0 < b; i0 += 4 * gsz) {
      int r[4], pr[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) r[j] = i0 + j * gsz < b ? a.rootc[i0 + j * gsz] : -1;
#pragma unroll
      for (int j = 0; j < 4; ++j) pr[j] = r[j] >= 0 ? gld(a.parent + r[j]) : -1;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (pr[j] != r[j]) a.rootc[i0 + j * gsz] = guf_find(a.parent, r[j]);
    }
    for (int i = b + gtid; i < e; i += gsz) {
      const int v = a.order[i];
      const int f = guf_find(a.parent, v);
      a.rootc[i] = f;
      atomicAdd(&a.size[f], 1);
      if (f == v) rn[atomicAdd(&nroots[(l + 1) % 3], 1)] = v;
    }
    grid_barrier(bar, fail, gen);
  }
  const int A = s_start[L];
  for (int i = gtid; i < A; i += gsz) a.enhn[a.order[i]] = a.acc[i];
  if (!ok) atomicOr(fail, 1);
  grid_barrier(bar, fail, gen);
  if (blockIdx.x == 0 && tid == 0 && a.status) a.status[0] = *fail ? kStatusNoProgress : L;
}

}  // namespace

hipError_t launch_tfce_coop(const TfceCoopArgs& a, hipStream_t stream) {
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tfce_coop_kernel, kCoopThreads, 0);
  if (e != hipSuccess) return e;
  int dev = 0, cus = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  // one workgroup per 4,096 elements, at most 24 and one per compute unit:
  // the grid barrier's cost grows with the workgroups it waits for (an
  // order-7 hemisphere took 2.74 ms with 24 workgroups, 2.81 ms with 32 and
  // 3.03 ms with 40; the 232k-voxel volume 6.50, 6.80 and 8.22 ms with 24,
  // 32 and 56; 160 took the hemisphere to 6.5 ms)
  int64_t want = std::min<int64_t>(24, ((int64_t)a.V + 4095) / 4096);
  int64_t cap = (int64_t)std::min(per_cu, 1) * cus;
  if (a.grid_limit > 0) cap = std::min<int64_t>(cap, a.grid_limit);
  const int grid = (int)std::max<int64_t>(1, std::min(want, cap));
  void* args[] = {const_cast<TfceCoopArgs*>(&a)};
  return hipLaunchCooperativeKernel(reinterpret_cast<void*>(tfce_coop_kernel), dim3(grid), dim3(kCoopThreads),
                                    args, 0, stream);
}

hipError_t launch_tfce_batch(const TfceBatchArgs& a, bool full, bool ring_only, int grid, int64_t n_global,
                             hipStream_t stream, int pair_grid) {
  hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(1), 0, stream, a.queue);
  hipLaunchKernelGGL(tfce_prepare_kernel, dim3((unsigned)a.n_items), dim3(kTfceThreads), 0, stream, a);
  if (n_global <= 0) return hipGetLastError();
  TfceBatchArgs g = a;
  g.n_items = n_global;
  if (pair_grid > 0)
    hipLaunchKernelGGL(tfce_pair_kernel, dim3(pair_grid), dim3(kPairThreads), 0, stream, g);
  else if (full && ring_only)
    hipLaunchKernelGGL((tfce_batch_kernel<true, true>), dim3(grid), dim3(kTfceThreads), 0, stream, g);
  else if (full)
    hipLaunchKernelGGL((tfce_batch_kernel<true, false>), dim3(grid), dim3(kTfceThreads), 0, stream, g);
  else if (ring_only)
    hipLaunchKernelGGL((tfce_batch_kernel<false, true>), dim3(grid), dim3(kTfceThreads), 0, stream, g);
  else
    hipLaunchKernelGGL((tfce_batch_kernel<false, false>), dim3(grid), dim3(kTfceThreads), 0, stream, g);
  return hipGetLastError();
}

hipError_t launch_tfce_lds(const TfceBatchArgs& a, int64_t item0, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(tfce_lds_kernel, dim3(grid), dim3(kLdsThreads), 0, stream, a, item0 / a.n_signs);
  return hipGetLastError();
}

hipError_t launch_shared_T0(const TfceBatchArgs& a, float* T0, int zero_in_image, hipStream_t stream) {
  hipLaunchKernelGGL(shared_T0_kernel, dim3((unsigned)(a.n_maps * a.n_signs)), dim3(kTfceThreads), 0, stream, a, T0,
                     zero_in_image);
  return hipGetLastError();
}

hipError_t launch_sizepow(double* out, int32_t n, float E, hipStream_t stream) {
  hipLaunchKernelGGL(sizepow_kernel, dim3((unsigned)((n + 256) / 256)), dim3(256), 0, stream, out, n, (double)E);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, int64_t n, float v, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, v);
  return hipGetLastError();
}

}  // namespace tfcehip
