// Internal (non-ABI) declarations shared by the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfcehip {

// Level tables live in LDS; the reference visits about 101 levels
// (fast_tfce.hpp:34-39 with deltaT = max/100), an explicit step may ask for more.
constexpr int kMaxLevels = 1024;
constexpr int32_t kNever = 0x7fffffff;
constexpr int kTfceThreads = 256;
constexpr int kEllMaxWidth = 32;  // longest neighbour list kept in the fixed-width table
constexpr int kNposMaxWidth = 16;
// Neighbour-position entries of a surface with a local adjacency table: the
// position in the low 24 bits (kPosNone: never active), the neighbour's
// adjacency byte within the row in the high 8.
constexpr int32_t kPosMask = 0xffffff;
constexpr int32_t kPosNone = 0xffffff;  // widest table whose per-image neighbour positions a slot holds

// Status codes written per work item (>= 0: number of levels visited).
constexpr int32_t kStatusDegenerate = -2;  // max >= minT with a step <= 0: the reference never returns
constexpr int32_t kStatusTooManyLevels = -3;
constexpr int32_t kStatusNoProgress = -4;  // a union-find join exceeded its retry bound

// Surface flags.
constexpr int32_t kSurfSymmetric = 1;     // every edge is listed from both ends
constexpr int32_t kSurfZeroInLevels = 2;  // the full image holds zeros outside this element set
constexpr int32_t kSurfZeroInScale = 4;   // the rescale max includes those zeros
// Symmetric surface whose rows hold at most kRingMaxDeg neighbours, put in
// ring order (a closed triangle mesh's vertex link), with fewer than kRingNone
// elements.  nadj then holds one 16-byte packed row per element: the
// neighbours' indices in 18 bits each (kRingNone past the row's end), a link
// byte (bit j < 5 set iff neighbours j and j+1 are adjacent, bit 5 iff the
// last and the first are) and the row length; a neighbour-position row packs
// positions the same way (kRingRowWords words).
constexpr int32_t kSurfRing = 8;
constexpr int kRingMaxDeg = 6;
constexpr int32_t kRingNone = 0x3ffff;  // 18-bit position of an inactive neighbour
constexpr int32_t kRingRowWords = 4;
// Symmetric surface with a wide table (rows of more than kNposMaxWidth, the
// 26-connected voxel grids) and full row adjacency (adjw): the sweep keeps one
// 16-byte row per active element of the positions it joins (kCandRowWords
// words, -1 past the last), found once per image in element order: every
// neighbour activated earlier in its level, and one neighbour per group of
// mutually adjacent neighbours activated at earlier levels.  An element with
// more than four such candidates holds -2 - (its element) in the first word
// and is joined from the table.
constexpr int32_t kCandRowWords = 4;
// Most distinct full-adjacency rows a surface keeps as a table of distinct
// rows with a 16-bit row id per element (8 MB of rows at width 32); above
// that every element keeps its own row.
constexpr int32_t kAdjDictMaxRows = 65535;
// Largest surface the LDS-resident sweep holds (12 bytes per element plus the
// level tables in 160 KB of LDS), and the most levels it takes (a step of
// max/100 gives at most 102).
constexpr int32_t kLdsMaxV = 13248;
constexpr int kLdsMaxLevels = 128;

struct SurfDev {
  const int32_t* row_ptr;  // V+1
  const int32_t* col;      // nnz (indices local to the surface)
  const int32_t* ell;      // [V][ell_width] neighbour lists padded with -1, or null
  int32_t ell_width;       // largest list length (0: use the CSR)
  const uint8_t* nadj;     // [V][ell_width]: bit k of entry j set iff neighbours j and 8*(j/8)+k are
                           // adjacent (groups of 8 entries); [V][16] packed rows with kSurfRing;
                           // null if not built
  const uint32_t* adjw;    // [V][ell_width] for rows wider than kNposMaxWidth: bit k of entry j set iff
                           // neighbours j and k are adjacent; null if not built
  const uint16_t* adjid;   // [V] with adjw: the row of adjw that holds element v's masks (adjw then
                           // holds each distinct row once), or null: row v
  const double* sizepow;   // [V+1] (double)s^E, null when E == 1
  int64_t offset;          // first element of this surface inside a map row
  int32_t V;
  int32_t flags;
  float H, E;
};

struct TfceBatchArgs {
  const SurfDev* surfs;    // device array [n_surf]
  int32_t n_surf;
  int32_t n_maps;
  int32_t n_signs;         // 1 or 2 (2 = +map then -map)
  float sign0;             // sign used when n_signs == 1
  const float* maps;       // map p starts at maps + p * map_stride
  int64_t map_stride;
  const float* ext_T0;     // optional [n_items]: first threshold (shared-threshold mode)
  const float* vdensity;   // optional per-element weights (map column indexing), FULL mode
  float min_t;
  float dh;                // 0: dh = fl(T0 / 100) as the reference
  int32_t full_scale;      // FULL mode: compute the rescaled maximum
  // outputs, one entry per item; item = (surface * n_maps + map) * n_signs + sign
  float* raw_max;          // max over elements of the TFCE sum (no rescale)
  float* scaled_max;       // max over elements of fl(fl(tfce * fl(smax/100)) * vd)
  int32_t out_mul, out_add;  // scaled_max index = item * out_mul + out_add
  float* own_max;          // max over this surface's elements of sign*map (NaN ignored)
  int32_t* status;
  float* enhn;             // FULL mode: accumulate target, same layout as maps (may be null)
  int32_t labels_level;    // FULL mode debug: component labels after this level
  int32_t* labels;         // [V] vertex id of a component representative, -1 inactive
  // scratch
  int32_t* scratch;
  int64_t slot_stride;     // int32 words per slot
  int32_t v_max;
  int32_t np_w;            // words per element of the slot's neighbour-position table (0: none)
  int32_t* queue;          // dynamic work counter, zeroed before launch
  int32_t* prep;           // [n_items][kPrepWords] per-item maximum and level tables
  int64_t n_items;
};

// Per-item record written by the prepare kernel (int32 words).
constexpr int kPrepL = 0;     // levels visited, or a negative status
constexpr int kPrepOwn = 1;   // max of sign*map over the surface (float bits)
constexpr int kPrepT0 = 2;    // first threshold (float bits)
constexpr int kPrepDT = 3;    // step (float bits)
constexpr int kPrepT = 4;     // kMaxLevels thresholds, then kMaxLevels values of T^H
constexpr int kPrepWords = kPrepT + 2 * kMaxLevels;

// One image over the whole GPU (FULL mode, symmetric adjacency), after the
// item's record has been written by launch_tfce_batch with n_global = 0.
struct TfceCoopArgs {
  const int32_t* row_ptr;  // V+1
  const int32_t* col;
  const double* sizepow;   // null when E == 1
  float E;
  int32_t V;
  const float* x;          // image [V]
  float* enhn;             // [V], accumulated
  const int32_t* rec;      // the item's record (kPrepWords)
  int32_t* lev;            // [V] scratch
  int32_t* order;          // [V]
  int32_t* parent;         // [V]
  int32_t* size;           // [V]
  int32_t* roots[2];       // [V] each
  int32_t* rootc;          // [V] final root of the element at each position
  float* acc;              // [V] TFCE value by position
  int32_t* ctl;            // [kCoopCtlWords], the barrier words zero at launch
  int32_t* status;         // [1] or null
  int32_t grid_limit;      // at most this many workgroups (0: one per compute unit)
};
// ctl layout: level starts [kMaxLevels + 1], cursors [kMaxLevels], root-list
// lengths [3], failure flag [1], grid barrier counter and generation [2]
constexpr int kCoopBarWord = 2 * kMaxLevels + 5;
constexpr int kCoopCtlWords = kCoopBarWord + 2;
hipError_t launch_tfce_coop(const TfceCoopArgs& a, hipStream_t stream);

// Clears the work counters, writes every item's record (a.n_items) and runs
// the global-memory sweep over items [0, n_global) (none when n_global <= 0).
// pair_grid > 0: the paired sweep (MAX mode, ring-packed surfaces, both signs
// of every map, items in (surface, map, sign) order, thresholds >= 0) with
// that many 512-thread workgroups, each with its own slot, instead.
hipError_t launch_tfce_batch(const TfceBatchArgs& a, bool full, bool ring_only, int grid, int64_t n_global,
                             hipStream_t stream, int pair_grid = 0);
// The LDS-resident MAX sweep over items [item0, a.n_items) (symmetric surfaces
// of at most kLdsMaxV elements with a neighbour table; item0 a multiple of
// a.n_signs: one workgroup takes both signs of a map); its work counter is the
// second 64-bit word of a.queue.  Run after launch_tfce_batch, which writes
// the items' records and clears the counters.  Only for the step max/100
// (dh == 0): an item with more than kLdsMaxLevels levels gets NaN and
// status kStatusTooManyLevels.
hipError_t launch_tfce_lds(const TfceBatchArgs& a, int64_t item0, int grid, hipStream_t stream);
hipError_t launch_sizepow(double* out, int32_t n, float E, hipStream_t stream);
hipError_t launch_shared_T0(const TfceBatchArgs& a, float* T0, int zero_in_image, hipStream_t stream);
hipError_t launch_fill_f32(float* p, int64_t n, float v, hipStream_t stream);

// Slot layout: 8 arrays of the largest surface's element count.
inline int64_t tfce_slot_words(int32_t v_max, int32_t np_w) {
  return (int64_t)(8 + np_w) * ((v_max + 63) & ~63);
}

}  // namespace tfcehip
