// dbscan_hip.hip -- gfx950 kernels and C-ABI launcher for 2-D grid DBSCAN.
//
// Labels are bit-identical to the reference CPU path
// (csrc/cpu/dbscan_grid.hpp:136-458).  Pipeline, all on one stream, no host
// synchronisation (the cell count stays on the device):
//
//   k_bounds     cell-coordinate bounds of the points (per-block partials) and
//                clearing of the look-back status words
//   k_hist       digit histograms of every radix pass, in one read of X
//   k_sort_pass  stable LSD radix pass (x8, passes beyond the key width exit
//                at once).  Sorts the points themselves, (x, y, index) in
//                struct-of-arrays form, by a compressed cell key that is
//                recomputed from (x, y) in every pass, so no key array exists.
//                Ranking is wave64 ballot matching into per-wave LDS counters;
//                tile offsets come from a decoupled look-back per digit.
//   k_cells      cell runs of the sorted points (look-back scan): point->cell,
//                cell start, cell key
//   k_ngh        per cell, the 5 window columns as cell-index ranges (cells of
//                a column are contiguous and ky-ordered), the cell's index in
//                the reference's cell order, union-find init
//   k_core       core marking: full cells wholesale, else window counting
//   k_union      lock-free union-find over core cells (CAS hooks the larger
//                reference index onto the smaller)
//   k_flatten    read-only root walk, root flags in reference order
//   k_rank       exclusive scan of root flags = dense cluster ids
//   k_cluster    cluster id per cell
//   k_label      core labels, border assignment with the reference tiebreak,
//                and the scatter back to input order, fused
//
// Internal cell order is (kx signed, ky signed).  The reference orders cells
// by the signed int64 key (u32(kx) << 32 | u32(ky)) (dbscan_grid.hpp:176-178),
// i.e. kx signed then ky as uint32; within one column that puts ky >= 0 before
// ky < 0.  k_ngh computes each cell's reference index (a rotation within its
// column) and union-find runs on reference indices, so roots are minimum
// reference indices and the rank of each root is the reference's cluster id
// (dbscan_grid.hpp:392-398).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "dbscan_hip.h"

// Workspace of dbscan2d_hip_noalloc: one array per GPU, placed by the code
// object loader, held until the process ends, shared by every call on that
// GPU.  In host code its name is only a handle, never a device address: the
// address comes from hipGetSymbolAddress, once per device (scratch_address).
static_assert(DBSCAN_HIP_NOALLOC_SCRATCH_BYTES % 256 == 0, "whole 256-byte blocks");
static __device__ __align__(256) unsigned char scratch[DBSCAN_HIP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int32_t kNoise = -1;
// Every label of a call whose device-side checks failed (Params::error != 0).
constexpr int32_t kPoison = INT32_MIN;
// Radix digits are at most r bits wide (r = the configuration's radix_bits,
// at most kRadixBitsMax).  A key of b bits is sorted in P = min(ceil(b / r),
// kMaxPasses) passes of ceil(b / P) bits each, both chosen on the device once
// the bounds are known (r = 9: 25-bit keys take 3 passes of 9 bits, 28-bit
// keys 4 passes of 7 bits).  In a batch
// every cloud takes the batch's largest P and splits its own key bits over
// those P passes.
constexpr int kRadixBitsMax = 9;
constexpr int kRadixMax = 1 << kRadixBitsMax;
constexpr int kMaxPasses = 8;           // 64-bit keys at most
constexpr int kSortThreads = 512;
// A radix pass ranks tiles of kSortThreads * items points, items being one of
// 8, 12, 16 (the configuration's sort_items); status words are sized for the
// smallest tile.
constexpr int kSortTileMin = kSortThreads * 8;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;
// The cell pass's tile (a batch's tiles: up to kCellTile points of one cloud).
constexpr int kCellItems = 8;
constexpr int kCellTile = kScanThreads * kCellItems;
static_assert(kCellTile >= kScanTile, "look-back words are sized for kScanTile tiles");
constexpr int kMaxReduceBlocks = 512;
// kTileOrder: the look-back kernels (k_sort_pass, k_cells, k_rank) use the
// workgroup id as the tile index.  Workgroups of a 1-D grid are dispatched in
// increasing id order, so a tile only ever waits on tiles that are already
// resident or finished.  (A contended atomic ticket costs ~10 ns per
// workgroup here, which is tens of microseconds for these grids.)  Every wait
// is bounded by kSpinLimit and reports kErrSortSpin / kErrScanSpin instead of
// hanging.
constexpr uint32_t kSpinLimit = 1u << 26;         // bounded waits; see ErrFlag
constexpr int kTailPass = 4;  // radix passes from here on: one walking launch (k_sort_tail)
constexpr uint64_t kCountMask = (1ull << 48) - 1;

// Thread t < kRadixMax of a sort workgroup owns digits [t*D, (t+1)*D).
constexpr int kDigitsPerThread = kRadixMax > kSortThreads ? kRadixMax / kSortThreads : 1;
static_assert(kDigitsPerThread * kSortThreads >= kRadixMax, "every digit has an owner");

// Error bits left in Params::error when a bounded wait gives up or an index
// check fails (never expected).  A call that ends with any bit set writes
// kPoison to every label; callers read the word with dbscan2d_hip_error_async
// or dbscan2d_hip_check.
enum ErrFlag : int32_t {
  kErrSortSpin = DBSCAN_HIP_DEV_SORT_WAIT,
  kErrScanSpin = DBSCAN_HIP_DEV_SCAN_WAIT,
  kErrUfLoop = DBSCAN_HIP_DEV_UF_LOOP,
  kErrRange = DBSCAN_HIP_DEV_RANGE  // a computed index fell outside its array (write skipped)
};

// First bytes of the workspace.  The first 8 ints are the `info` block of
// dbscan_hip_layout.
struct Params {
  int32_t kxmin, kxmax, kymin, kymax;
  int32_t num_cells;
  int32_t npasses;
  int32_t error;
  int32_t col_total;  // batches: column-index entries in use
  int32_t num_roots;  // roots over all cells (k_rank): bounds every cluster id
  int32_t pad_[3];
  // Everything below is cleared by k_bounds at the start of every call, in
  // 16-byte stores from here to the end of the look-back status words.
  uint32_t tile_counter[16];
  uint32_t hist[kMaxPasses][kRadixMax];
};
constexpr size_t kParamsZeroBegin = offsetof(Params, tile_counter);
static_assert(kParamsZeroBegin % 16 == 0, "the cleared range starts 16-byte aligned");

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

// (int32)floor(v), with out-of-range and NaN mapped to INT32_MIN -- what the
// reference's x86 build produces for dbscan_grid.hpp:162-163.
__device__ __forceinline__ int32_t cell_coord(float v) {
  const float f = floorf(v);
  return (f >= -2147483648.0f && f < 2147483648.0f) ? static_cast<int32_t>(f) : INT32_MIN;
}

struct KeyMap {
  uint32_t kxmin, kymin;  // subtracted (mod 2^32) from kx, ky
  uint32_t bx, by;        // bit widths of the x and y parts
  int npasses;            // radix passes covering bx + by bits (>= 1)
  uint32_t dbits;         // digit width of passes 1, 2, ...
  uint32_t dlow;          // digit width of pass 0: the key bits the others leave
};

__device__ __forceinline__ uint32_t bit_width(uint32_t v) {
  return v ? 32u - static_cast<uint32_t>(__clz(static_cast<int>(v))) : 0u;
}

// passes > 0: split the key over that many passes (the count chosen when the
// bounds were reduced, or a batch's common count).  Otherwise the passes are
// as few as digits of at most `rbits` bits allow, and never more than
// kMaxPasses (a wider key then takes wider digits, up to kRadixBitsMax).
__device__ __forceinline__ KeyMap make_keymap(int32_t kxmin, int32_t kxmax, int32_t kymin,
                                              int32_t kymax, int passes = 0,
                                              int rbits = kRadixBitsMax) {
  KeyMap m;
  m.kxmin = static_cast<uint32_t>(kxmin);
  m.kymin = static_cast<uint32_t>(kymin);
  m.bx = bit_width(static_cast<uint32_t>(kxmax) - m.kxmin);
  m.by = bit_width(static_cast<uint32_t>(kymax) - m.kymin);
  const int nbits = static_cast<int>(m.bx + m.by);
  m.npasses = passes > 0 ? passes
                         : (nbits ? min((nbits + rbits - 1) / rbits, kMaxPasses) : 1);
  m.dbits = nbits ? static_cast<uint32_t>((nbits + m.npasses - 1) / m.npasses) : 1u;
  // The passes' spare bits go to pass 0 (a 26-bit key in 3 passes: 8 + 9 + 9):
  // its digits, the lowest key bits, are spread the most evenly over a tile,
  // so fewer and longer runs help it most.  With fewer key bits than passes
  // every pass keeps dbits and the top digits are 0.
  const int rest = (m.npasses - 1) * static_cast<int>(m.dbits);
  m.dlow = nbits > rest ? static_cast<uint32_t>(nbits - rest) : m.dbits;
  return m;
}

__device__ __forceinline__ KeyMap load_keymap(const Params* p) {
  return make_keymap(p->kxmin, p->kxmax, p->kymin, p->kymax, p->npasses);
}

// One cloud of a batch (dbscan2d_hip_batched): its point range in the packed
// input, its scalar parameters (computed on the host exactly as for a single
// call), and what the kernels derive for it.  A single call is not a batch
// and keeps its parameters in kernel arguments and Params.
struct CloudInfo {
  int32_t begin, end;            // points [begin, end) of the packed input
  float inv_r, eps_sq, best_init;
  int32_t ms;                    // narrowed min_samples; INT32_MAX: every point is noise
  int32_t kxmin, kxmax, kymin, kymax;  // cell bounds (k_bounds_seg, atomics)
  int32_t cell_begin, cell_end;  // its cells (k_cells)
  int32_t col_base;              // first column-index entry, or -1 (binary search)
  int32_t pad[3];
};
static_assert(sizeof(CloudInfo) == 64, "CloudInfo is 64 bytes");

// A tile of consecutive points of one cloud.  `first` is the global index of
// the cloud's first tile (look-backs stop there).
struct SegTile {
  int32_t cloud, start, count, first;
};

__device__ __forceinline__ KeyMap cloud_keymap(const CloudInfo& ci, int passes,
                                               int rbits = kRadixBitsMax) {
  return make_keymap(ci.kxmin, ci.kxmax, ci.kymin, ci.kymax, passes, rbits);
}

// Scalar parameters of the cloud a cell belongs to.  The per-cell kernels
// take them from one of two sources: OneCloud (a single call: kernel
// arguments) or ManyClouds (a batch: the cell's cloud record).
struct CellParams {
  float eps_sq, best_init;
  int32_t ms;
};
struct OneCloud {
  static constexpr bool kUniform = true;  // every cell has the same parameters
  float eps_sq, best_init;
  int32_t ms;
  __device__ __forceinline__ CellParams at(int32_t) const { return {eps_sq, best_init, ms}; }
};
struct ManyClouds {
  static constexpr bool kUniform = false;
  const int32_t* __restrict__ cell_cloud;
  const CloudInfo* __restrict__ clouds;
  __device__ __forceinline__ CellParams at(int32_t c) const {
    const CloudInfo& ci = clouds[cell_cloud[c]];
    return {ci.eps_sq, ci.best_init, ci.ms};
  }
};

// Column index: for x' < 2^bx <= max_cols, col_range[x'] = (first cell, end
// cell) of column x' in cell order (both 0 for an empty column).
__device__ __forceinline__ bool has_col_index(const KeyMap& m, int64_t max_cols) {
  return m.bx < 32 && (1ll << m.bx) <= max_cols;
}

// Relative cell coordinates (x', y') = (kx - kxmin, ky - kymin) of a point.
__device__ __forceinline__ void rel_cell(float x, float y, float inv_r, const KeyMap& m,
                                         uint32_t& xr, uint32_t& yr) {
  xr = static_cast<uint32_t>(cell_coord(x * inv_r)) - m.kxmin;
  yr = static_cast<uint32_t>(cell_coord(y * inv_r)) - m.kymin;
}

// Compressed sort key: x' above y', bx + by bits in all.
__device__ __forceinline__ uint64_t sort_key(float x, float y, float inv_r, const KeyMap& m) {
  uint32_t xr, yr;
  rel_cell(x, y, inv_r, m, xr, yr);
  return (static_cast<uint64_t>(xr) << m.by) | yr;
}

// Width and digit of radix pass `pass`: pass 0 takes the lowest dlow bits,
// pass p > 0 the dbits above dlow + (p - 1) dbits.
__device__ __forceinline__ uint32_t digit_width(int pass, const KeyMap& m) {
  return pass == 0 ? m.dlow : m.dbits;
}
__device__ __forceinline__ uint32_t key_digit(uint64_t key, int pass, const KeyMap& m) {
  const uint32_t shift = pass == 0 ? 0u : m.dlow + m.dbits * static_cast<uint32_t>(pass - 1);
  return static_cast<uint32_t>(key >> shift) & ((1u << digit_width(pass, m)) - 1u);
}

// Cell key in internal order: (x' << 32 | y').
__device__ __forceinline__ uint64_t cell_key64(uint32_t xr, uint32_t yr) {
  return (static_cast<uint64_t>(xr) << 32) | yr;
}

// Relaxed agent-scope atomic load and store: coherent across the device's
// L2s, but ordering nothing else.  Every word published through them is
// self-contained (a look-back status word carries its tag and its count
// together; a union-find parent is the whole datum), so readers need no
// ordering.  Where data other than the word itself must become visible (the
// sort tail's pass hand-off, the last workgroup of k_small), the code fences
// explicitly with __threadfence() before the store and after the load.
template <typename T>
__device__ __forceinline__ T ld_relaxed(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_relaxed(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive scan over a block of NT threads; `s_w` holds NT/64 words.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total, uint32_t* s_w) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) {
    const uint32_t s = s_w[i];
    off += (i < w) ? s : 0u;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + inc - v;
}

__device__ __forceinline__ int lowest_lane(uint64_t m) {
  return __ffsll(static_cast<unsigned long long>(m)) - 1;
}

// Decoupled look-back for a single running total per tile (tags: 1 = tile
// aggregate, 2 = inclusive prefix), by the 64 lanes of one wave: a round trip
// reads the status words of up to 64 preceding tiles at once (lane k reads
// tile j - k), adds the aggregates up to the first inclusive prefix, and
// stops there; a word not yet published stops the round short of it and is
// waited for, within kSpinLimit.  The first round, and every round after one
// that found the nearest word unpublished, reads that one word only: with
// thousands of tiles resident, 64-word polls of words that are not there yet
// were a large share of the memory traffic (k_cells 71 -> 61 us at 10M
// points).  Returns the exclusive prefix of `tile` (in every lane) and
// publishes the tile's inclusive prefix.
__device__ uint64_t lookback_wave(uint64_t* status, uint32_t tile, uint64_t aggregate,
                                  int32_t* error) {
  constexpr uint64_t kAgg = 1ull << 48, kPre = 2ull << 48;
  const int lane = threadIdx.x & 63;
  if (tile == 0) {
    if (lane == 0) st_relaxed(status, kPre | aggregate);
    return 0;
  }
  if (lane == 0) st_relaxed(status + tile, kAgg | aggregate);
  uint32_t excl = 0;  // counts are below 2^31
  int64_t j = static_cast<int64_t>(tile) - 1;
  uint32_t spins = 0;
  int width = 1;  // words read per round: 64 while words are being consumed
  for (;;) {
    const int64_t q = j - lane;
    const uint64_t s = lane >= width ? 0ull
                       : q >= 0      ? ld_relaxed(status + q)
                                     : kPre;  // before tile 0: prefix 0
    const uint64_t tag = s >> 48;
    const uint64_t pre = __ballot(tag == 2), miss = __ballot(tag == 0);
    const int first_pre = pre ? lowest_lane(pre) : 64;
    const int first_miss = miss ? lowest_lane(miss) : 64;
    const bool done = first_pre < first_miss;  // a prefix before any gap
    const int use = done ? first_pre + 1 : first_miss;
    uint32_t v = lane < use ? static_cast<uint32_t>(s & kCountMask) : 0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    excl += v;
    if (done) break;
    j -= use;
    width = use == 0 ? 1 : 64;
    if (use == 0) {
      if (++spins > kSpinLimit) {
        if (lane == 0) atomicOr(error, static_cast<int32_t>(kErrScanSpin));
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  if (lane == 0) st_relaxed(status + tile, kPre | (static_cast<uint64_t>(excl) + aggregate));
  return excl;
}

// Union-find over reference cell indices.  Parents only ever move to smaller
// indices (dbscan_grid.hpp:80-117), so every walk ends at the component's
// minimum index.  Loads are agent-scope atomics so that writes by other
// workgroups (on other XCDs) are seen within the kernel.
// Read-only root walk, bounded by kSpinLimit steps (a longer walk can only
// come from corrupted parents: kErrUfLoop, and the walk stops there).
__device__ __forceinline__ int32_t uf_find_ro(const int32_t* uf, int32_t x, int32_t* error) {
  int32_t p = ld_relaxed(uf + x);
  for (uint32_t it = 0; p != x; ++it) {
    if (it >= kSpinLimit) {
      atomicOr(error, static_cast<int32_t>(kErrUfLoop));
      break;
    }
    x = p;
    p = ld_relaxed(uf + x);
  }
  return x;
}

__device__ __forceinline__ int32_t uf_find_halving(int32_t* uf, int32_t x) {
  int32_t p = ld_relaxed(uf + x);
  while (p != x) {
    const int32_t gp = ld_relaxed(uf + p);
    if (gp != p) st_relaxed(uf + x, gp);
    x = p;
    p = gp;
  }
  return x;
}

__device__ void uf_unite(int32_t* uf, int32_t a, int32_t b, int32_t* error) {
  for (uint32_t it = 0; it < kSpinLimit; ++it) {
    a = uf_find_halving(uf, a);
    b = uf_find_halving(uf, b);
    if (a == b) return;
    const int32_t hi = a > b ? a : b, lo = a > b ? b : a;
    if (atomicCAS(uf + hi, hi, lo) == hi) return;
  }
  atomicOr(error, static_cast<int32_t>(kErrUfLoop));
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_fill_i32(int32_t* __restrict__ out, int64_t n,
                                                  int32_t v, int32_t* __restrict__ error_out) {
  if (error_out && blockIdx.x == 0 && threadIdx.x == 0) error_out[0] = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * 256)
    out[i] = v;
}

template <int kWaves = 4>
__device__ __forceinline__ void block_minmax4(int32_t v[4], int32_t (*s)[4]) {
  // v = {min, max, min, max}
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    v[0] = min(v[0], __shfl_xor(v[0], d, 64));
    v[1] = max(v[1], __shfl_xor(v[1], d, 64));
    v[2] = min(v[2], __shfl_xor(v[2], d, 64));
    v[3] = max(v[3], __shfl_xor(v[3], d, 64));
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) s[w][k] = v[k];
  }
  __syncthreads();
  v[0] = s[0][0]; v[1] = s[0][1]; v[2] = s[0][2]; v[3] = s[0][3];
#pragma unroll
  for (int i = 1; i < kWaves; ++i) {
    v[0] = min(v[0], s[i][0]);
    v[1] = max(v[1], s[i][1]);
    v[2] = min(v[2], s[i][2]);
    v[3] = max(v[3], s[i][3]);
  }
  __syncthreads();
}

// Cell-coordinate bounds, one partial per block; also clears the words the
// later kernels accumulate into (histograms, tile counters, look-back status).
__global__ __launch_bounds__(256) void k_bounds(const float2* __restrict__ X, int64_t n,
                                                float inv_r, int4* __restrict__ partials,
                                                uint4* __restrict__ zero, int64_t zero16) {
  const int64_t gtid = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
  for (int64_t z = gtid; z < zero16; z += stride) zero[z] = make_uint4(0, 0, 0, 0);

  int32_t v[4] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  auto take = [&](float x, float y) {
    const int32_t kx = cell_coord(x * inv_r), ky = cell_coord(y * inv_r);
    v[0] = min(v[0], kx); v[1] = max(v[1], kx);
    v[2] = min(v[2], ky); v[3] = max(v[3], ky);
  };
  if (reinterpret_cast<uintptr_t>(X) & 15u) {  // only 8-byte aligned: one point per load
    for (int64_t i = gtid; i < n; i += stride) take(X[i].x, X[i].y);
  } else {
    // two points per 16-byte load, two loads in flight per thread
    const float4* __restrict__ X4 = reinterpret_cast<const float4*>(X);
    const int64_t pairs = n >> 1;
    int64_t q = gtid;
    for (; q + stride < pairs; q += 2 * stride) {
      const float4 a = X4[q], b = X4[q + stride];
      take(a.x, a.y); take(a.z, a.w); take(b.x, b.y); take(b.z, b.w);
    }
    if (q < pairs) {
      const float4 a = X4[q];
      take(a.x, a.y); take(a.z, a.w);
    }
    if ((n & 1) && gtid == 0) take(X[n - 1].x, X[n - 1].y);
  }
  __shared__ int32_t s[4][4];
  block_minmax4(v, s);
  if (threadIdx.x == 0) partials[blockIdx.x] = make_int4(v[0], v[1], v[2], v[3]);
}

// Reduces the bound partials, publishes them, and accumulates the digit
// histogram of every radix pass.  One large workgroup per CU: each adds its
// histogram into the global one at the end, and every such atomic lands on
// one of only npasses x 128 words (512 workgroups of 256: 36 us for 10M
// points; 256 of 1,024: 30 us; 128 or fewer leave CUs idle).
constexpr int kHistThreads = 1024;
constexpr int kHistBlocks = 256;
__global__ __launch_bounds__(kHistThreads) void k_hist(const float2* __restrict__ X, int64_t n,
                                              float inv_r, const int4* __restrict__ partials,
                                              int nparts, Params* __restrict__ prm,
                                              int2* __restrict__ col_range, int64_t max_cols,
                                              int rbits, int32_t error0) {
  constexpr int T = kHistThreads;
  __shared__ int32_t s_red[T / 64][4];
  __shared__ uint32_t s_hist[kMaxPasses][kRadixMax];
  int32_t v[4] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  for (int i = threadIdx.x; i < nparts; i += T) {
    const int4 q = partials[i];
    v[0] = min(v[0], q.x); v[1] = max(v[1], q.y);
    v[2] = min(v[2], q.z); v[3] = max(v[3], q.w);
  }
  block_minmax4<T / 64>(v, s_red);
  const KeyMap m = make_keymap(v[0], v[1], v[2], v[3], 0, rbits);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    prm->kxmin = v[0]; prm->kxmax = v[1]; prm->kymin = v[2]; prm->kymax = v[3];
    prm->npasses = m.npasses;
    prm->num_cells = 0;
    prm->error = error0;
    prm->col_total = 0;
  }
  for (int e = threadIdx.x; e < kMaxPasses * kRadixMax; e += T) (&s_hist[0][0])[e] = 0u;
  __syncthreads();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * T;
  const int64_t gtid = static_cast<int64_t>(blockIdx.x) * T + threadIdx.x;
  if (has_col_index(m, max_cols))
    for (int64_t q = gtid; q < (1ll << m.bx); q += stride) col_range[q] = make_int2(0, 0);
  auto take = [&](float x, float y) {
    const uint64_t key = sort_key(x, y, inv_r, m);
    for (int ps = 0; ps < m.npasses; ++ps) atomicAdd(&s_hist[ps][key_digit(key, ps, m)], 1u);
  };
  if (reinterpret_cast<uintptr_t>(X) & 15u) {  // only 8-byte aligned: one point per load
    for (int64_t i = gtid; i < n; i += stride) take(X[i].x, X[i].y);
  } else {
    // two points per 16-byte load, two loads in flight per thread
    const float4* __restrict__ X4 = reinterpret_cast<const float4*>(X);
    const int64_t pairs = n >> 1;
    int64_t q = gtid;
    for (; q + stride < pairs; q += 2 * stride) {
      const float4 a = X4[q], b = X4[q + stride];
      take(a.x, a.y); take(a.z, a.w); take(b.x, b.y); take(b.z, b.w);
    }
    if (q < pairs) {
      const float4 a = X4[q];
      take(a.x, a.y); take(a.z, a.w);
    }
    if ((n & 1) && gtid == 0) take(X[n - 1].x, X[n - 1].y);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m.npasses * kRadixMax; e += T) {
    const uint32_t c = (&s_hist[0][0])[e];
    if (c) atomicAdd(&prm->hist[0][0] + e, c);
  }
}

// ---- Batches: per-cloud bounds, key maps and histograms -------------------

// Cell-coordinate bounds of each cloud: one block per tile, merged into the
// cloud's record with atomics (the record holds INT_MAX / INT_MIN on entry).
// Also clears the words the later kernels accumulate into.
__global__ __launch_bounds__(256) void k_bounds_seg(const float2* __restrict__ X,
                                                    const SegTile* __restrict__ tiles,
                                                    CloudInfo* __restrict__ clouds,
                                                    uint4* __restrict__ zero, int64_t zero16) {
  const int64_t gtid = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
  for (int64_t z = gtid; z < zero16; z += stride) zero[z] = make_uint4(0, 0, 0, 0);
  const SegTile st = tiles[blockIdx.x];
  const float inv_r = clouds[st.cloud].inv_r;
  int32_t v[4] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  for (int i = threadIdx.x; i < st.count; i += 256) {
    const float2 p = X[st.start + i];
    const int32_t kx = cell_coord(p.x * inv_r), ky = cell_coord(p.y * inv_r);
    v[0] = min(v[0], kx); v[1] = max(v[1], kx);
    v[2] = min(v[2], ky); v[3] = max(v[3], ky);
  }
  __shared__ int32_t s[4][4];
  block_minmax4(v, s);
  if (threadIdx.x == 0) {
    CloudInfo* ci = clouds + st.cloud;
    atomicMin(&ci->kxmin, v[0]);
    atomicMax(&ci->kxmax, v[1]);
    atomicMin(&ci->kymin, v[2]);
    atomicMax(&ci->kymax, v[3]);
  }
}

// One block: the batch's pass count (the largest any cloud needs) and each
// cloud's column-index base (an exclusive scan of 2^bx over the clouds; a
// cloud whose columns would pass max_cols searches instead).
__global__ __launch_bounds__(1024) void k_keys_seg(Params* __restrict__ prm,
                                                   CloudInfo* __restrict__ clouds,
                                                   int32_t num_clouds, int64_t max_cols,
                                                   int rbits, int32_t error0) {
  __shared__ int32_t s_max[16];
  __shared__ uint32_t s_scan[16];
  __shared__ int64_t s_run;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int32_t pmax = 1;
  for (int32_t i = t; i < num_clouds; i += 1024) {
    const CloudInfo& ci = clouds[i];
    if (ci.end > ci.begin) pmax = max(pmax, cloud_keymap(ci, 0, rbits).npasses);
  }
  for (int d = 32; d >= 1; d >>= 1) pmax = max(pmax, __shfl_xor(pmax, d, 64));
  if (lane == 0) s_max[w] = pmax;
  if (t == 0) s_run = 0;
  __syncthreads();
  if (t == 0) {
    int32_t m = 1;
    for (int i = 0; i < 16; ++i) m = max(m, s_max[i]);
    prm->npasses = m;
    prm->num_cells = 0;
    prm->error = error0;
  }
  for (int32_t b0 = 0; b0 < num_clouds; b0 += 1024) {
    const int32_t i = b0 + t;
    uint32_t cols = 0;
    if (i < num_clouds) {
      const CloudInfo& ci = clouds[i];
      const uint32_t bx = cloud_keymap(ci, 1).bx;
      // a cloud of more than 2^20 columns always searches
      if (ci.end > ci.begin && bx <= 20) cols = 1u << bx;
    }
    uint32_t tot;
    const uint32_t off = block_excl_scan<1024>(cols, tot, s_scan);
    const int64_t at = s_run + off;
    if (i < num_clouds)
      clouds[i].col_base = (cols && at + cols <= max_cols) ? static_cast<int32_t>(at) : -1;
    __syncthreads();
    if (t == 0) s_run += tot;
    __syncthreads();
  }
  if (t == 0) prm->col_total = static_cast<int32_t>(min<int64_t>(s_run, max_cols));
}

// Digit histograms of every pass, per cloud.  Block b takes a contiguous run
// of tiles and accumulates in LDS while the cloud stays the same, flushing the
// non-zero counts to the cloud's histogram when it changes.  Also clears the
// column index.
__global__ __launch_bounds__(256) void k_hist_seg(const float2* __restrict__ X,
                                                  const SegTile* __restrict__ tiles,
                                                  int32_t num_tiles, int32_t tiles_per_block,
                                                  const CloudInfo* __restrict__ clouds,
                                                  const Params* __restrict__ prm,
                                                  uint32_t* __restrict__ cloud_hist,
                                                  int2* __restrict__ col_range) {
  __shared__ uint32_t s_hist[kMaxPasses][kRadixMax];
  const int t = threadIdx.x;
  const int P = prm->npasses;
  const int64_t gtid = static_cast<int64_t>(blockIdx.x) * 256 + t;
  for (int64_t q = gtid; q < prm->col_total; q += static_cast<int64_t>(gridDim.x) * 256)
    col_range[q] = make_int2(0, 0);
  for (int e = t; e < kMaxPasses * kRadixMax; e += 256) (&s_hist[0][0])[e] = 0u;
  const int32_t t0 = blockIdx.x * tiles_per_block;
  const int32_t t1 = min(num_tiles, t0 + tiles_per_block);
  int32_t cur = -1;
  auto flush = [&]() {
    __syncthreads();
    if (cur >= 0) {
      uint32_t* h = cloud_hist + static_cast<size_t>(cur) * kMaxPasses * kRadixMax;
      for (int e = t; e < P * kRadixMax; e += 256) {
        const uint32_t c = (&s_hist[0][0])[e];
        if (c) {
          atomicAdd(h + e, c);
          (&s_hist[0][0])[e] = 0u;
        }
      }
    }
    __syncthreads();
  };
  for (int32_t ti = t0; ti < t1; ++ti) {
    const SegTile st = tiles[ti];
    if (st.cloud != cur) {
      flush();
      cur = st.cloud;
    }
    const CloudInfo& ci = clouds[st.cloud];
    const KeyMap m = cloud_keymap(ci, P);
    for (int i = t; i < st.count; i += 256) {
      const float2 p = X[st.start + i];
      const uint64_t key = sort_key(p.x, p.y, ci.inv_r, m);
      for (int ps = 0; ps < P; ++ps) atomicAdd(&s_hist[ps][key_digit(key, ps, m)], 1u);
    }
  }
  flush();
}

// Predecessor status words a radix-pass look-back reads per round trip.
constexpr int kSortLookback = 4;

// Sum of the earlier tiles' counts of one digit (decoupled look-back).  The
// status words of kSortLookback predecessors are read per round trip; a word tagged
// `tag_pre` is an inclusive prefix (stop), `tag_agg` a tile's own count
// (add, keep going), anything else is not yet published (or left by an
// earlier pass) and is waited for, within kSpinLimit.
// Tiles below `first` (another cloud of a batch) are not part of the sum.
__device__ __forceinline__ uint64_t digit_lookback(const uint64_t* __restrict__ status,
                                                   int64_t tile, int64_t first, int digit,
                                                   uint64_t tag_agg, uint64_t tag_pre,
                                                   int32_t* error, int stride = kRadixMax) {
  uint64_t excl = 0;
  int64_t j = tile - 1;
  uint32_t spins = 0;
  while (j >= first) {
    uint64_t w[kSortLookback];
#pragma unroll
    for (int k = 0; k < kSortLookback; ++k)
      w[k] = (j - k >= first)
                 ? ld_relaxed(status + static_cast<uint64_t>(j - k) * stride + digit)
                 : tag_pre;
    // consume the published words in order: `used` of them, `stop` at a prefix
    int used = 0;
    bool stop = false, blocked = false;
#pragma unroll
    for (int k = 0; k < kSortLookback; ++k) {
      const uint64_t tag = w[k] & ~kCountMask;
      const bool ready = tag == tag_agg || tag == tag_pre;
      if (!stop && !blocked) {
        if (ready) {
          excl += w[k] & kCountMask;
          ++used;
          stop = tag == tag_pre;
        } else {
          blocked = true;  // not yet published
        }
      }
    }
    if (stop) break;
    j -= used;
    if (used == 0) {
      if (++spins > kSpinLimit) {
        atomicOr(error, static_cast<int32_t>(kErrSortSpin));
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  return excl;
}

// One stable LSD radix pass over digit `pass` of the compressed cell key.
// kFirst: read the interleaved input X and emit original indices.
// A tile of kSortThreads * kItems points is ranked in LDS (wave64 ballot matching into
// per-wave counters, item by item, which keeps equal digits in input order),
// offset by the global digit histogram and a look-back over earlier tiles,
// then reordered by digit through one LDS buffer, array by array, so that runs
// of equal digits leave as contiguous writes.  Thread t owns digits
// [t*kDigitsPerThread, (t+1)*kDigitsPerThread).
// Every pass but the last writes (x, y) as 8-byte pairs over the first 8n
// bytes of its x buffer (the y array follows the x array in the workspace,
// so the pairs fit) and the next pass reads them back: one store per point
// instead of two, and fewer partly written cache lines at the ends of the
// digit runs.  The last pass writes the x and y arrays the cell stages read.
// kSeg (batches): tile t is tiles[t], a run of one
// cloud's points; the cloud's own histogram, key map and scale apply, the
// look-back stops at the cloud's first tile, and the cloud's points stay in
// its own range [begin, end), so every cloud is sorted on its own.
// One tile, `tile`, of pass `pass` (pass < npasses); the workgroup's last
// LDS access is followed by a barrier, so a caller may run it in a loop.
template <bool kFirst, bool kSeg, int kItems>
__device__ __forceinline__ void sort_pass_tile(
    const float2* __restrict__ X, const float* __restrict__ in_x, const float* __restrict__ in_y,
    const uint32_t* __restrict__ in_i, float* __restrict__ out_x, float* __restrict__ out_y,
    uint32_t* __restrict__ out_i, Params* __restrict__ prm, uint64_t* __restrict__ status,
    int pass, int64_t n64, float inv_r, const SegTile* __restrict__ tiles,
    const CloudInfo* __restrict__ clouds, const uint32_t* __restrict__ cloud_hist,
    uint32_t tile) {
  constexpr int kTile = kSortThreads * kItems;
  const uint32_t n = static_cast<uint32_t>(n64);  // n < 2^31: 32-bit offsets
  {
    KeyMap m;
    uint32_t base, tile_n, out_base = 0, first = 0;
    const uint32_t* __restrict__ hist;
    if constexpr (kSeg) {
      const SegTile st = tiles[tile];
      const CloudInfo& ci = clouds[st.cloud];
      m = cloud_keymap(ci, prm->npasses);
      base = static_cast<uint32_t>(st.start);
      tile_n = static_cast<uint32_t>(st.count);
      out_base = static_cast<uint32_t>(ci.begin);
      first = static_cast<uint32_t>(st.first);
      inv_r = ci.inv_r;
      hist = cloud_hist + (static_cast<size_t>(st.cloud) * kMaxPasses + pass) * kRadixMax;
    } else {
      m = load_keymap(prm);
      base = tile * kTile;
      tile_n = min(static_cast<uint32_t>(kTile), n - base);
      hist = prm->hist[pass];
    }
    constexpr int D = kDigitsPerThread;

    __shared__ uint64_t s_buf[kTile];  // (x, y) pairs, then indices in its first half
    __shared__ uint16_t s_d[kTile];
    __shared__ uint32_t s_wcnt[kSortThreads / 64][kRadixMax];
    __shared__ uint32_t s_goff[kRadixMax];
    __shared__ uint32_t s_dstart[kRadixMax];
    // the two block scans finish (a barrier follows) before s_d is first
    // written, and s_d is last read before the tile's closing barrier: the
    // scan words live in its first bytes (12-point tiles of 9-bit digits then
    // fit 80 KB, two workgroups per CU)
    uint32_t* s_scan = reinterpret_cast<uint32_t*>(s_d);
    static_assert(kTile * sizeof(uint16_t) >= (kSortThreads / 64) * sizeof(uint32_t), "");

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // the pass's digits: only these are counted, published and looked back on
    const uint32_t dw = digit_width(pass, m);
    const int nd = 1 << dw;
    for (int e = t; e < (kSortThreads / 64) * nd; e += kSortThreads)
      s_wcnt[e >> dw][e & (nd - 1)] = 0u;
    // global start of each owned digit: exclusive scan of the pass histogram
    uint32_t goff[D];
    {
      uint32_t h[D], sum = 0;
#pragma unroll
      for (int q = 0; q < D; ++q) {
        h[q] = t * D + q < kRadixMax ? hist[t * D + q] : 0u;
        sum += h[q];
      }
      uint32_t tot;
      uint32_t run = out_base + block_excl_scan<kSortThreads>(sum, tot, s_scan);
#pragma unroll
      for (int q = 0; q < D; ++q) {
        goff[q] = run;
        run += h[q];
      }
    }
    // local index of item k of this lane: wl + k * 64
    const uint32_t wl = static_cast<uint32_t>(wave) * (64 * kItems) + lane;

    float x[kItems], y[kItems];
    uint32_t id[kItems], dg[kItems], rk[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      // unconditional loads (a clamped index past the end) keep the per-item
      // values in plain registers; validity is tested where it matters
      const uint32_t l = wl + k * 64;
      const uint32_t e = base + l;
      const uint32_t ec = l < tile_n ? e : base + tile_n - 1;
      if (kFirst) {
        const float2 p = X[ec];
        x[k] = p.x; y[k] = p.y; id[k] = e;
      } else {
        const float2 p = reinterpret_cast<const float2*>(in_x)[ec];  // an earlier pass's pairs
        x[k] = p.x; y[k] = p.y; id[k] = in_i[ec];
      }
    }
    // all loads of the tile are in flight before the first key is computed
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < kItems; ++k) dg[k] = key_digit(sort_key(x[k], y[k], inv_r, m), pass, m);

    // Rank within the wave's segment, item by item (stable): lanes holding the
    // same digit are found with one ballot per digit bit; the lowest such lane
    // bumps the wave's counter for that digit.
    const uint64_t lt_mask = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      const bool valid = wl + k * 64 < tile_n;
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < kRadixBitsMax; ++b) {
        if (static_cast<uint32_t>(b) < dw) {
          const bool bit = (dg[k] >> b) & 1u;
          const uint64_t bal = __ballot(bit);
          peers &= bit ? bal : ~bal;
        }
      }
      const uint32_t below = static_cast<uint32_t>(__popcll(peers & lt_mask));
      uint32_t old = 0;
      if (valid) old = s_wcnt[wave][dg[k]];
      if (valid && below == 0) s_wcnt[wave][dg[k]] = old + static_cast<uint32_t>(__popcll(peers));
      rk[k] = old + below;
    }
    __syncthreads();

    // Owned digits: wave offsets, tile totals, aggregate publication.
    const uint64_t tag_agg = static_cast<uint64_t>(2 * pass + 1) << 48;
    const uint64_t tag_pre = static_cast<uint64_t>(2 * pass + 2) << 48;
    uint32_t dtot[D], dsum = 0;
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const int d = t * D + q;
      dtot[q] = 0;
      if (d < nd) {
        // per-wave counts -> exclusive offsets of the waves
#pragma unroll
        for (int w = 0; w < kSortThreads / 64; ++w) {
          const uint32_t c = s_wcnt[w][d];
          s_wcnt[w][d] = dtot[q];
          dtot[q] += c;
        }
        st_relaxed(status + static_cast<uint64_t>(tile) * kRadixMax + d,
                   (tile == first ? tag_pre : tag_agg) | dtot[q]);
      }
      dsum += dtot[q];
    }
    uint32_t ttot;
    uint32_t dstart = block_excl_scan<kSortThreads>(dsum, ttot, s_scan);
    uint32_t dst0[D];  // tile-local start of each owned digit
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const int d = t * D + q;
      dst0[q] = dstart;
      if (d < nd) s_dstart[d] = dstart;
      dstart += dtot[q];
    }
    __syncthreads();

    // Tile-local position of every item (kept in rk); the digits, and (x, y),
    // go to LDS in position order before the look-back, so that x, y and the
    // digits are no longer held in registers while it waits.
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      rk[k] += s_dstart[dg[k]] + s_wcnt[wave][dg[k]];
      if (wl + k * 64 < tile_n) {
        s_d[rk[k]] = static_cast<uint16_t>(dg[k]);
        s_buf[rk[k]] = (static_cast<uint64_t>(__float_as_uint(y[k])) << 32) | __float_as_uint(x[k]);
      }
    }
    // Global start of each owned digit: the earlier tiles' counts by look-back.
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const int d = t * D + q;
      if (d >= nd) break;
      uint64_t excl = 0;
      if (tile > first) {
        excl = digit_lookback(status, static_cast<int64_t>(tile), static_cast<int64_t>(first), d,
                              tag_agg, tag_pre, &prm->error);
        st_relaxed(status + static_cast<uint64_t>(tile) * kRadixMax + d, tag_pre | (excl + dtot[q]));
      }
      s_goff[d] = goff[q] + static_cast<uint32_t>(excl) - dst0[q];
    }
    __syncthreads();
    // The output position of each element this thread writes (element j = t +
    // kSortThreads q of the tile, in digit order), then (x, y) and the index
    // through the one LDS buffer.
    uint32_t pos[kItems];
#pragma unroll
    for (int q = 0; q < kItems; ++q) {
      const int j = t + q * kSortThreads;
      pos[q] = j < static_cast<int>(tile_n) ? s_goff[s_d[j]] + static_cast<uint32_t>(j) : n;
    }
    if (pass + 1 == prm->npasses) {
#pragma unroll
      for (int q = 0; q < kItems; ++q) {
        const int j = t + q * kSortThreads;
        if (pos[q] < n) {
          const uint64_t v = s_buf[j];
          out_x[pos[q]] = __uint_as_float(static_cast<uint32_t>(v));
          out_y[pos[q]] = __uint_as_float(static_cast<uint32_t>(v >> 32));
        } else if (j < static_cast<int>(tile_n)) {
          atomicOr(&prm->error, static_cast<int32_t>(kErrRange));
        }
      }
    } else {
      uint64_t* __restrict__ out_xy = reinterpret_cast<uint64_t*>(out_x);
#pragma unroll
      for (int q = 0; q < kItems; ++q) {
        const int j = t + q * kSortThreads;
        if (pos[q] < n) {
          out_xy[pos[q]] = s_buf[j];
        } else if (j < static_cast<int>(tile_n)) {
          atomicOr(&prm->error, static_cast<int32_t>(kErrRange));
        }
      }
    }
    __syncthreads();
    uint32_t* s_idx = reinterpret_cast<uint32_t*>(s_buf);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      if (wl + k * 64 < tile_n) s_idx[rk[k]] = id[k];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kItems; ++q) {
      if (pos[q] < n) out_i[pos[q]] = s_idx[t + q * kSortThreads];
    }
    __syncthreads();
  }
}

// One radix pass, one tile per workgroup.  Tiles of up to 12 points per
// thread fit two workgroups per CU in LDS; the register budget is held to
// match (4 waves per SIMD: at most 128 VGPRs).
template <bool kFirst, bool kSeg, int kItems>
__global__ __launch_bounds__(kSortThreads, kItems <= 12 ? 4 : 2) void k_sort_pass(
    const float2* __restrict__ X, const float* __restrict__ in_x, const float* __restrict__ in_y,
    const uint32_t* __restrict__ in_i, float* __restrict__ out_x, float* __restrict__ out_y,
    uint32_t* __restrict__ out_i, Params* __restrict__ prm, uint64_t* __restrict__ status,
    int pass, int64_t n64, float inv_r, const SegTile* __restrict__ tiles,
    const CloudInfo* __restrict__ clouds, const uint32_t* __restrict__ cloud_hist,
    uint32_t num_tiles) {
  if (pass >= prm->npasses || blockIdx.x >= num_tiles) return;
  sort_pass_tile<kFirst, kSeg, kItems>(X, in_x, in_y, in_i, out_x, out_y, out_i, prm, status,
                                       pass, n64, inv_r, tiles, clouds, cloud_hist,
                                       blockIdx.x);  // see kTileOrder
}

// The radix passes from kTailPass on (needed only for keys wider than four
// digits; for the usual key the launch returns at once), in one launch that
// needs no workgroup to be resident beside another: a workgroup claims work
// items by an atomic ticket, item k being tile k mod num_tiles of pass
// kTailPass + k / num_tiles, so every tile of a pass is claimed after all
// tiles of the passes before it and after the earlier tiles of its own pass.
// A tile's look-back waits only on earlier tiles of its pass, claimed by
// workgroups that are running; a pass's first tiles wait until every tile of
// the pass before has finished (a done counter per pass, agent-scope fence
// before the count and after the wait, which also invalidates this CU's
// vector cache before the earlier pass's output is read).  A workgroup holds
// one tile at a time and finishes it before it claims the next, so the
// claimed tiles always make progress whatever the residency.  Every wait is
// bounded by kSpinLimit.  Pass p reads buffer set a if p is odd, else b.
// tile_counter[0] is the ticket, tile_counter[1 + q] the tiles done of pass
// first_pass + q (cleared by the call's first kernel).  first_pass is
// kTailPass, or 2 for inputs of few tiles (kTailFromOneTiles), whose passes
// cost more in kernel boundaries than in work.
template <bool kSeg, int kItems>
__global__ __launch_bounds__(kSortThreads) void k_sort_tail(
    float* __restrict__ xs_a, float* __restrict__ ys_a, uint32_t* __restrict__ idx_a,
    float* __restrict__ xs_b, float* __restrict__ ys_b, uint32_t* __restrict__ idx_b,
    Params* __restrict__ prm, uint64_t* __restrict__ status, int64_t n64, float inv_r,
    const SegTile* __restrict__ tiles, const CloudInfo* __restrict__ clouds,
    const uint32_t* __restrict__ cloud_hist, int first_pass, uint32_t num_tiles) {
  static_assert(kMaxPasses <= 16, "one done counter per tail pass");
  const int np = prm->npasses;
  if (np <= first_pass || num_tiles == 0) return;
  const uint32_t items = static_cast<uint32_t>(np - first_pass) * num_tiles;
  __shared__ uint32_t s_ticket;
  // Every branch around the barriers below is wave-uniform: wave 0 as a
  // whole claims the ticket, waits and counts (lane 0 contributes the only
  // increment), the others skip those blocks.  A branch on threadIdx.x == 0
  // at the end of the body and another at its start were merged by the
  // compiler across the loop's back edge into a divergent exit, which let
  // the other lanes run into the next iteration's barrier ahead of thread 0.
  const bool wave0 = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x)) == 0;
  const uint32_t one = threadIdx.x == 0 ? 1u : 0u;
  for (;;) {
    if (wave0) {
      const uint32_t old = atomicAdd(&prm->tile_counter[0], one);
      s_ticket = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(old), 0));
    }
    __syncthreads();
    const uint32_t k = static_cast<uint32_t>(
        __builtin_amdgcn_readfirstlane(static_cast<int>(s_ticket)));
    __syncthreads();
    if (k >= items) break;
    const int q = static_cast<int>(k / num_tiles);
    const int pass = first_pass + q;
    const uint32_t tile = k - static_cast<uint32_t>(q) * num_tiles;
    if (q > 0) {
      if (wave0) {
        uint32_t spins = 0;
        while (static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(
                   ld_relaxed(&prm->tile_counter[q])))) < num_tiles) {
          if (++spins > kSpinLimit) {
            atomicOr(&prm->error, static_cast<int32_t>(kErrSortSpin));
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      __syncthreads();
      __threadfence();  // acquire in every wave before the earlier pass's output is read
    }
    const bool from_a = (pass & 1) != 0;
    sort_pass_tile<false, kSeg, kItems>(
        nullptr, from_a ? xs_a : xs_b, from_a ? ys_a : ys_b, from_a ? idx_a : idx_b,
        from_a ? xs_b : xs_a, from_a ? ys_b : ys_a, from_a ? idx_b : idx_a, prm, status, pass,
        n64, inv_r, tiles, clouds, cloud_hist, tile);
    // every thread's writes of the tile are made visible before it counts
    // as done
    __threadfence();
    __syncthreads();
    if (wave0) atomicAdd(&prm->tile_counter[q + 1], one);
  }
}

// Cell runs over the sorted points (dbscan_grid.hpp:206-227): flags where the
// key changes, block scan plus look-back, then point->cell, cell starts and
// cell keys.  Blocked layout: thread t owns kCellItems consecutive points,
// read and written with 16-byte vector accesses where aligned.
// kSeg (batches): tile t is tiles[t], a run of one
// cloud's points; a cloud's first point always opens a cell, keys are the
// cloud's own, and the cell numbering runs on across clouds.  Each cloud's
// cell range, and each cell's cloud, are recorded.
template <bool kSeg>
__global__ __launch_bounds__(kScanThreads) void k_cells(
    const float* __restrict__ xs_a, const float* __restrict__ ys_a,
    const float* __restrict__ xs_b, const float* __restrict__ ys_b, Params* __restrict__ prm,
    uint64_t* __restrict__ status, int32_t* __restrict__ point_cell,
    int32_t* __restrict__ cell_start, uint64_t* __restrict__ cell_key,
    uint32_t* __restrict__ cell_flags, uint4* __restrict__ cell_box, int2* __restrict__ col_range,
    int64_t max_cols, int64_t n64, float inv_r, const SegTile* __restrict__ tiles,
    CloudInfo* __restrict__ clouds, int32_t* __restrict__ cell_cloud) {
  static_assert(kCellItems % 4 == 0, "whole float4s per thread");
  __shared__ uint64_t s_last[kScanThreads];
  __shared__ uint32_t s_scan[kScanThreads / 64];
  __shared__ uint64_t s_prefix;
  const uint32_t n = static_cast<uint32_t>(n64);  // n < 2^31
  const bool in_a = (prm->npasses & 1) != 0;
  const float* __restrict__ xs = in_a ? xs_a : xs_b;
  const float* __restrict__ ys = in_a ? ys_a : ys_b;
  const int t = threadIdx.x;
  const uint32_t tile = blockIdx.x;  // see kTileOrder
  KeyMap m;
  uint32_t base, end, cbeg, cend;  // tile [base, end) of the cloud's [cbeg, cend)
  bool colidx;
  int64_t col_base = 0;
  int32_t cloud = 0;
  if constexpr (kSeg) {
    const SegTile st = tiles[tile];
    cloud = st.cloud;
    const CloudInfo& ci = clouds[cloud];
    m = cloud_keymap(ci, prm->npasses);
    inv_r = ci.inv_r;
    base = static_cast<uint32_t>(st.start);
    end = base + static_cast<uint32_t>(st.count);
    cbeg = static_cast<uint32_t>(ci.begin);
    cend = static_cast<uint32_t>(ci.end);
    colidx = ci.col_base >= 0;
    col_base = ci.col_base;
  } else {
    m = load_keymap(prm);
    base = tile * kCellTile;
    end = min(base + kCellTile, n);
    cbeg = 0;
    cend = n;
    colidx = has_col_index(m, max_cols);
  }
  const uint32_t e0 = base + static_cast<uint32_t>(t) * kCellItems;
  const bool whole = e0 + kCellItems <= end;

  // thread 0's previous point (the last of the tile before), loaded with the
  // tile's own points rather than after them
  float bx = 0.f, by = 0.f;
  if (t == 0 && base > cbeg) {
    bx = xs[base - 1];
    by = ys[base - 1];
  }
  float px[kCellItems], py[kCellItems];
  if (whole && (e0 & 3u) == 0) {
#pragma unroll
    for (int v = 0; v < kCellItems / 4; ++v) {
      const float4 xv = *reinterpret_cast<const float4*>(xs + e0 + 4 * v);
      const float4 yv = *reinterpret_cast<const float4*>(ys + e0 + 4 * v);
      px[4 * v] = xv.x; px[4 * v + 1] = xv.y; px[4 * v + 2] = xv.z; px[4 * v + 3] = xv.w;
      py[4 * v] = yv.x; py[4 * v + 1] = yv.y; py[4 * v + 2] = yv.z; py[4 * v + 3] = yv.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kCellItems; ++j) {
      const uint32_t ec = e0 + j < end ? e0 + j : end - 1;  // clamped: defined, never used
      px[j] = xs[ec];
      py[j] = ys[ec];
    }
  }
  uint64_t key[kCellItems];
#pragma unroll
  for (int j = 0; j < kCellItems; ++j) {
    uint32_t xr, yr;
    rel_cell(px[j], py[j], inv_r, m, xr, yr);
    key[j] = cell_key64(xr, yr);
  }
  s_last[t] = key[kCellItems - 1];
  __syncthreads();
  uint64_t prev = 0;
  bool have_prev = false;  // the previous point exists and is of the same cloud
  if (t > 0) {
    prev = s_last[t - 1];
    have_prev = true;
  } else if (base > cbeg) {
    uint32_t xr, yr;
    rel_cell(bx, by, inv_r, m, xr, yr);
    prev = cell_key64(xr, yr);
    have_prev = true;
  }
  uint32_t flags = 0, newcol = 0, hadprev = 0, cnt = 0;
  uint32_t prev_x[kCellItems];
#pragma unroll
  for (int j = 0; j < kCellItems; ++j) {
    const bool f = e0 + j < end && (!have_prev || key[j] != prev);
    flags |= static_cast<uint32_t>(f) << j;
    newcol |= static_cast<uint32_t>(f && (!have_prev || (key[j] >> 32) != (prev >> 32))) << j;
    hadprev |= static_cast<uint32_t>(have_prev) << j;
    prev_x[j] = have_prev ? static_cast<uint32_t>(prev >> 32) : 0u;
    cnt += f;
    prev = key[j];
    have_prev = true;
  }
  uint32_t ttot;
  const uint32_t texcl = block_excl_scan<kScanThreads>(cnt, ttot, s_scan);
  if (t < 64) {
    const uint64_t pre = lookback_wave(status, tile, ttot, &prm->error);
    if (t == 0) s_prefix = pre;
  }
  __syncthreads();
  int32_t c = static_cast<int32_t>(s_prefix + texcl) - 1;
  int32_t pc[kCellItems];
#pragma unroll
  for (int j = 0; j < kCellItems; ++j) {
    if ((flags >> j) & 1u) {
      ++c;
      if (static_cast<uint32_t>(c) < n) {
        cell_start[c] = static_cast<int32_t>(e0 + j);
        cell_key[c] = key[j];
        cell_flags[c] = 0u;
        cell_box[c] = make_uint4(0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u);  // empty box for k_bbox
        if constexpr (kSeg) {
          cell_cloud[c] = cloud;
          if (e0 + j == cbeg) clouds[cloud].cell_begin = c;
        }
        if (colidx && ((newcol >> j) & 1u)) {
          col_range[col_base + static_cast<int64_t>(key[j] >> 32)].x = c;
          if ((hadprev >> j) & 1u) col_range[col_base + prev_x[j]].y = c;
        }
      } else {
        atomicOr(&prm->error, static_cast<int32_t>(kErrRange));
      }
    }
    pc[j] = c;
  }
  if (whole && (e0 & 3u) == 0) {
#pragma unroll
    for (int v = 0; v < kCellItems / 4; ++v)
      *reinterpret_cast<int4*>(point_cell + e0 + 4 * v) =
          make_int4(pc[4 * v], pc[4 * v + 1], pc[4 * v + 2], pc[4 * v + 3]);
  } else {
#pragma unroll
    for (int j = 0; j < kCellItems; ++j)
      if (e0 + j < end) point_cell[e0 + j] = pc[j];
  }
  if (e0 <= cend - 1 && cend - 1 < e0 + kCellItems && cend - 1 < end) {
    // this thread holds the cloud's last point
    const int j = static_cast<int>((cend - 1) - e0);
    const int32_t last = pc[j];
    if constexpr (kSeg) clouds[cloud].cell_end = last + 1;
    if (colidx) col_range[col_base + static_cast<int64_t>(key[j] >> 32)].y = last + 1;
    if (cend == n) {
      prm->num_cells = last + 1;
      cell_start[last + 1] = static_cast<int32_t>(n);
    }
  }
}

__device__ __forceinline__ int32_t lower_bound_u64(const uint64_t* __restrict__ a, int32_t cnt,
                                                   uint64_t v) {
  int32_t lo = 0, hi = cnt;
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The x array of the radix buffer pair that the last pass did not write: free
// from the cell pass until the border pass; k_ngh clears it and the 3x3 pair
// pass records its hits there (a 7-bit slot mask per cell).
__device__ __forceinline__ uint32_t* spare_words(const Params* prm, float* xs_a, float* xs_b) {
  return reinterpret_cast<uint32_t*>((prm->npasses & 1) ? xs_b : xs_a);
}

// Window columns, reference index, union-find init.  For column x' + dx the
// cells with y' in [y'-2, y'+2] are a contiguous run [j0, j1) of cells, so
// the 25-cell window of the reference (dbscan_grid.hpp:239-252) is 5 runs,
// and their order (dx-major, then ky ascending) is the reference's window
// offset order off = (dx+2)*5 + (dy+2).  With a column index the searches
// run inside one column; otherwise over all of the cloud's cells.
// kSeg (batches): a cell's cloud supplies its cell range, bounds and column
// index; reference indices stay inside the cloud's cell range.
template <bool kSeg>
__device__ __forceinline__ void ngh_cells(Params* __restrict__ prm,
                                          const uint64_t* __restrict__ cell_key,
                                          const int2* __restrict__ col_range, int64_t max_cols,
                                          int32_t* __restrict__ ngh,
                                          int32_t* __restrict__ cell_ref,
                                          int32_t* __restrict__ uf,
                                          const int32_t* __restrict__ cell_cloud,
                                          const CloudInfo* __restrict__ clouds,
                                          float* __restrict__ xs_a, float* __restrict__ xs_b,
                                          uint32_t block, uint32_t nblocks) {
  const int32_t C = prm->num_cells;
  uint32_t* __restrict__ hits = spare_words(prm, xs_a, xs_b);
  for (int64_t c64 = static_cast<int64_t>(block) * 256 + threadIdx.x; c64 < C;
       c64 += static_cast<int64_t>(nblocks) * 256) {
    const int32_t c = static_cast<int32_t>(c64);
    int32_t cb, ce, kymin, kymax;  // the cloud's cells [cb, ce) and y bounds
    bool colidx;
    int64_t col_base, ncols;
    if constexpr (kSeg) {
      const CloudInfo& ci = clouds[cell_cloud[c]];
      cb = ci.cell_begin;
      ce = ci.cell_end;
      kymin = ci.kymin;
      kymax = ci.kymax;
      colidx = ci.col_base >= 0;
      col_base = ci.col_base;
      ncols = colidx ? (1ll << cloud_keymap(ci, 1).bx) : 0;
    } else {
      const KeyMap m = load_keymap(prm);
      cb = 0;
      ce = C;
      kymin = prm->kymin;
      kymax = prm->kymax;
      colidx = has_col_index(m, max_cols);
      col_base = 0;
      ncols = colidx ? (1ll << m.bx) : 0;
    }
    // cells of column x (x < 2^32): [lo, hi)
    auto column = [&](int64_t x, int32_t& lo, int32_t& hi) {
      if (colidx) {
        if (x >= ncols) {
          lo = hi = 0;
        } else {
          const int2 r = col_range[col_base + x];
          lo = r.x;
          hi = r.y;
        }
      } else {
        lo = cb + lower_bound_u64(cell_key + cb, ce - cb, static_cast<uint64_t>(x) << 32);
        hi = x == 0xFFFFFFFFll
                 ? ce
                 : cb + lower_bound_u64(cell_key + cb, ce - cb, static_cast<uint64_t>(x + 1) << 32);
      }
    };
    const bool straddle = kymin < 0 && kymax >= 0;
    const uint32_t yr0 = 0u - static_cast<uint32_t>(kymin);  // y' of ky == 0
    const uint64_t key = cell_key[c];
    const uint32_t xr = static_cast<uint32_t>(key >> 32), yr = static_cast<uint32_t>(key);
    const uint32_t ylo = yr >= 2u ? yr - 2u : 0u;
    const uint32_t yhi = yr <= 0xFFFFFFFDu ? yr + 2u : 0xFFFFFFFFu;
    int32_t* row = ngh + c64 * 10;
    // Every key compared below lies in a known column (inside that column's
    // [lo, hi)), so only its y' word (the low half) is loaded.
    const uint32_t* __restrict__ key_y = reinterpret_cast<const uint32_t*>(cell_key);
    static_assert(sizeof(uint64_t) == 2 * sizeof(uint32_t), "");
    // The five columns' searches run in lock-step, so that each step is one
    // round of five independent loads rather than five dependent ones.
    int32_t lo[5], hi[5], base[5], len[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const int64_t cx = static_cast<int64_t>(xr) + (k - 2);
      lo[k] = hi[k] = 0;
      if (cx >= 0 && cx <= 0xFFFFFFFFll) column(cx, lo[k], hi[k]);
      base[k] = lo[k];
      len[k] = hi[k] - lo[k];
    }
    const int32_t own_lo = lo[2], own_hi = hi[2];
    // lower bound of (column, ylo): first cell of the column at y' >= ylo
    // (all five loads of a step issued before any is used: a finished or
    // empty search loads cell c, which exists)
    for (uint32_t it = 0; it < 40; ++it) {  // len < 2^31: at most 32 halvings
      uint32_t v[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) v[k] = key_y[2 * (len[k] > 0 ? base[k] + (len[k] >> 1) : c)];
      bool more = false;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int32_t half = len[k] >> 1;
        const bool right = v[k] < ylo;
        if (len[k] > 0) {
          base[k] = right ? base[k] + half + 1 : base[k];
          len[k] = right ? len[k] - half - 1 : half;
        }
        more |= len[k] > 0;
      }
      if (!more) break;
    }
    // the column's cells up to y' <= yhi: at most 5 (one per y'), keys
    // ascending, so the qualifying ones are a prefix
    uint32_t w[5][5];
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
      for (int q = 0; q < 5; ++q) w[k][q] = key_y[2 * (base[k] + q < hi[k] ? base[k] + q : c)];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      int32_t cnt = 0;
#pragma unroll
      for (int q = 0; q < 5; ++q) cnt += (base[k] + q < hi[k] && w[k][q] <= yhi) ? 1 : 0;
      row[2 * k] = base[k];
      row[2 * k + 1] = base[k] + cnt;
    }
    int32_t r = c;
    if (straddle) {
      const int32_t split = own_lo + lower_bound_u64(cell_key + own_lo, own_hi - own_lo,
                                                     (static_cast<uint64_t>(xr) << 32) | yr0);
      r = (yr >= yr0) ? own_lo + (c - split) : c + (own_hi - split);
    }
    if (r < cb || r >= ce) {
      atomicOr(&prm->error, static_cast<int32_t>(kErrRange));
      r = c;
    }
    cell_ref[c] = r;
    uf[r] = r;
    hits[c] = 0u;
  }
}

template <bool kSeg>
__global__ __launch_bounds__(256) void k_ngh(Params* __restrict__ prm,
                                             const uint64_t* __restrict__ cell_key,
                                             const int2* __restrict__ col_range, int64_t max_cols,
                                             int32_t* __restrict__ ngh,
                                             int32_t* __restrict__ cell_ref,
                                             int32_t* __restrict__ uf,
                                             const int32_t* __restrict__ cell_cloud,
                                             const CloudInfo* __restrict__ clouds,
                                             float* __restrict__ xs_a, float* __restrict__ xs_b) {
  ngh_cells<kSeg>(prm, cell_key, col_range, max_cols, ngh, cell_ref, uf, cell_cloud, clouds,
                  xs_a, xs_b, blockIdx.x, gridDim.x);
}

// Ordered-integer image of a float (monotone for non-NaN values), so that
// per-cell bounds can use integer atomicMin / atomicMax.
__device__ __forceinline__ uint32_t f2ord(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Cell bounding box [xmin, xmax, ymin, ymax] decoded from its ordered form.
__device__ __forceinline__ float4 load_box(const uint4* __restrict__ cell_box, int32_t c) {
  const uint4 u = cell_box[c];
  return make_float4(ord2f(u.x), ord2f(u.y), ord2f(u.z), ord2f(u.w));
}

// Squared distance from a point to a box, as the reference prunes with it
// (dbscan_grid.hpp:119-127): by monotone rounding it never exceeds the
// computed d2 of any point inside the box, so "> eps_sq" proves that no point
// of the cell is within eps.  fmaxf drops NaN terms, which only weakens the
// prune.
__device__ __forceinline__ float point_box_d2(float x, float y, float4 b) {
  const float dx = fmaxf(fmaxf(b.x - x, x - b.y), 0.0f);
  const float dy = fmaxf(fmaxf(b.z - y, y - b.w), 0.0f);
  return dx * dx + dy * dy;
}

// dbscan_grid.hpp:129-134
__device__ __forceinline__ float box_box_d2(float4 a, float4 b) {
  const float dx = fmaxf(fmaxf(b.x - a.y, a.x - b.y), 0.0f);
  const float dy = fmaxf(fmaxf(b.z - a.w, a.z - b.w), 0.0f);
  return dx * dx + dy * dy;
}

__device__ __forceinline__ int32_t bcast(int32_t v, int src) {
  return __builtin_amdgcn_readlane(v, src);
}
__device__ __forceinline__ float bcastf(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// Per-cell bounding boxes of the non-NaN points (the reference's per-cell
// boxes, dbscan_grid.hpp:332-349, used only for sound prunes), and core
// marking pass 1 (dbscan_grid.hpp:270-279): every point of a cell with at
// least min_samples points is core.  Each thread takes kBoxItems consecutive
// sorted points and reduces their runs of equal cells in registers; a run
// wholly inside the thread is a whole cell, stored directly.  The runs that
// touch a thread's ends are joined across the wave by a segmented suffix
// reduction over the threads' head runs (shuffles, once per thread rather
// than per point); a cell that stays inside the wave is stored directly, one
// reaching past it waits in the wave's LDS slot, and the slots of the
// workgroup are merged before they are folded, with atomics, into the
// sentinel written by k_cells (a cell of 300k points then costs 4 atomics per
// 2,048 points).
// cell_flags bits: kCellHasNaN and, for full cells, kCellHasCore are set by
// k_bbox; k_core_sparse sets the core bits of the other cells.
constexpr uint32_t kCellHasCore = 1u, kCellHasNonCore = 2u, kCellHasNaN = 4u;

constexpr int kBoxItems = 8;
constexpr int kBoxTile = 256 * kBoxItems;

// A run's box (ordered-integer bounds) and flag bits.
struct BoxPiece {
  uint32_t mnx, mxx, mny, mxy, fl;
};
__device__ __forceinline__ BoxPiece piece_empty() {
  return BoxPiece{0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u, 0u};
}
__device__ __forceinline__ void piece_add(BoxPiece& a, const BoxPiece& b) {
  a.mnx = min(a.mnx, b.mnx);
  a.mxx = max(a.mxx, b.mxx);
  a.mny = min(a.mny, b.mny);
  a.mxy = max(a.mxy, b.mxy);
  a.fl |= b.fl;
}
__device__ __forceinline__ BoxPiece piece_shfl_down(const BoxPiece& p, int d) {
  auto f = [&](uint32_t v) {
    return static_cast<uint32_t>(__shfl_down(static_cast<int>(v), d, 64));
  };
  return BoxPiece{f(p.mnx), f(p.mxx), f(p.mny), f(p.mxy), f(p.fl)};
}
// A whole cell: its box stored over the sentinel.
__device__ __forceinline__ void piece_store(uint4* __restrict__ cell_box,
                                            uint32_t* __restrict__ cell_flags, int32_t c,
                                            const BoxPiece& p) {
  cell_box[c] = make_uint4(p.mnx, p.mxx, p.mny, p.mxy);
  if (p.fl) atomicOr(cell_flags + c, p.fl);
}
// Part of a cell: folded in with atomics.
__device__ __forceinline__ void piece_fold(uint4* __restrict__ cell_box,
                                           uint32_t* __restrict__ cell_flags, int32_t c,
                                           const BoxPiece& p) {
  if (p.mnx != 0xFFFFFFFFu) {
    atomicMin(&cell_box[c].x, p.mnx);
    atomicMax(&cell_box[c].y, p.mxx);
    atomicMin(&cell_box[c].z, p.mny);
    atomicMax(&cell_box[c].w, p.mxy);
  }
  if (p.fl) atomicOr(cell_flags + c, p.fl);
}

template <class PP>
__device__ __forceinline__ void bbox_tile(const Params* __restrict__ prm,
                                          const float* __restrict__ xs_a,
                                          const float* __restrict__ ys_a,
                                          const float* __restrict__ xs_b,
                                          const float* __restrict__ ys_b,
                                          const int32_t* __restrict__ point_cell,
                                          const int32_t* __restrict__ cell_start,
                                          uint4* __restrict__ cell_box,
                                          uint32_t* __restrict__ cell_flags,
                                          uint8_t* __restrict__ is_core, int64_t n, PP pp,
                                          uint32_t tile) {
  // per wave: [2w] the cell continuing in from before the wave, [2w + 1] the
  // cell continuing past it (-1: none)
  __shared__ int32_t s_key[8];
  __shared__ BoxPiece s_val[8];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const bool in_a = (prm->npasses & 1) != 0;
  const float* __restrict__ xs = in_a ? xs_a : xs_b;
  const float* __restrict__ ys = in_a ? ys_a : ys_b;
  const int64_t e0 = static_cast<int64_t>(tile) * kBoxTile + static_cast<int64_t>(t) * kBoxItems;
  int32_t pc[kBoxItems];
  float px[kBoxItems], py[kBoxItems];
  if (e0 + kBoxItems <= n) {
    const int4 c0 = *reinterpret_cast<const int4*>(point_cell + e0);
    const int4 c1 = *reinterpret_cast<const int4*>(point_cell + e0 + 4);
    const float4 x0 = *reinterpret_cast<const float4*>(xs + e0);
    const float4 x1 = *reinterpret_cast<const float4*>(xs + e0 + 4);
    const float4 y0 = *reinterpret_cast<const float4*>(ys + e0);
    const float4 y1 = *reinterpret_cast<const float4*>(ys + e0 + 4);
    pc[0] = c0.x; pc[1] = c0.y; pc[2] = c0.z; pc[3] = c0.w;
    pc[4] = c1.x; pc[5] = c1.y; pc[6] = c1.z; pc[7] = c1.w;
    px[0] = x0.x; px[1] = x0.y; px[2] = x0.z; px[3] = x0.w;
    px[4] = x1.x; px[5] = x1.y; px[6] = x1.z; px[7] = x1.w;
    py[0] = y0.x; py[1] = y0.y; py[2] = y0.z; py[3] = y0.w;
    py[4] = y1.x; py[5] = y1.y; py[6] = y1.z; py[7] = y1.w;
  } else {
#pragma unroll
    for (int j = 0; j < kBoxItems; ++j) {
      const bool v = e0 + j < n;
      pc[j] = v ? point_cell[e0 + j] : -1;  // -1: past n, never stored
      px[j] = v ? xs[e0 + j] : 0.f;
      py[j] = v ? ys[e0 + j] : 0.f;
    }
  }
  // full cells: every point core (the loads of all items issued together)
  bool full[kBoxItems];
#pragma unroll
  for (int j = 0; j < kBoxItems; ++j) {
    const int32_t c = pc[j] >= 0 ? pc[j] : 0;
    full[j] = pc[j] >= 0 && cell_start[c + 1] - cell_start[c] >= pp.at(c).ms;
  }
  if (e0 + kBoxItems <= n) {
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      lo |= static_cast<uint32_t>(full[j]) << (8 * j);
      hi |= static_cast<uint32_t>(full[j + 4]) << (8 * j);
    }
    *reinterpret_cast<uint2*>(is_core + e0) = make_uint2(lo, hi);  // other cells: k_core_sparse
  } else {
#pragma unroll
    for (int j = 0; j < kBoxItems; ++j)
      if (e0 + j < n) is_core[e0 + j] = full[j] ? 1 : 0;
  }
  auto item = [&](int j) {
    BoxPiece p = piece_empty();
    if (!isnan(px[j]) && !isnan(py[j])) {
      p.mnx = p.mxx = f2ord(px[j]);
      p.mny = p.mxy = f2ord(py[j]);
    } else {
      // the box ignores NaN points; its x maximum becomes +inf, so that no
      // box test takes the cell as lying within eps (the other tests only
      // weaken: a bigger box is never farther than a point of it)
      p.mxx = f2ord(__builtin_inff());
      p.fl = kCellHasNaN;
    }
    if (full[j]) p.fl |= kCellHasCore;
    return p;
  };
  // the thread's runs: head (holds item 0), tail (holds the last item); runs
  // strictly between them are whole cells
  BoxPiece head = piece_empty(), cur = item(0);
  int32_t curc = pc[0];
  int run_start = 0;
#pragma unroll
  for (int j = 1; j < kBoxItems; ++j) {
    if (pc[j] == curc) {
      piece_add(cur, item(j));
    } else {
      if (run_start == 0) head = cur;
      else if (curc >= 0) piece_store(cell_box, cell_flags, curc, cur);
      curc = pc[j];
      cur = item(j);
      run_start = j;
    }
  }
  const BoxPiece tail = cur;
  const bool single = run_start == 0;
  if (single) head = cur;
  const int32_t hc = pc[0], tc = pc[kBoxItems - 1];
  int32_t prev_tc = __shfl_up(tc, 1, 64), next_hc = __shfl_down(hc, 1, 64);
  if (lane == 0) prev_tc = (e0 > 0 && e0 - 1 < n) ? point_cell[e0 - 1] : -1;
  if (lane == 63) next_hc = (e0 + kBoxItems < n) ? point_cell[e0 + kBoxItems] : -1;
  const bool link_in = hc >= 0 && prev_tc == hc;
  const bool link_out = tc >= 0 && next_hc == tc;
  // whole cells among the end runs
  if (single) {
    if (!link_in && !link_out && hc >= 0) piece_store(cell_box, cell_flags, hc, head);
  } else {
    if (!link_in && hc >= 0) piece_store(cell_box, cell_flags, hc, head);
    if (!link_out && tc >= 0) piece_store(cell_box, cell_flags, tc, tail);
  }
  // S = head runs reduced over the following threads of the same cell (a
  // thread's head cell recurs in the next thread only if the thread is one run)
  BoxPiece S = head;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t o = __shfl_down(hc, d, 64);
    const BoxPiece q = piece_shfl_down(S, d);
    if (lane + d < 64 && o == hc) piece_add(S, q);
  }
  // the cell that runs out of the wave through lane 63 (-2: none)
  const int32_t pass63 = bcast(single && link_out ? hc : -2, 63);
  const BoxPiece S1 = piece_shfl_down(S, 1);
  if (lane == 0) s_key[2 * wv] = s_key[2 * wv + 1] = -1;
  __syncthreads();
  // a cell that starts in this thread's tail run and continues into the next
  if (link_out && (!single || !link_in)) {
    BoxPiece total = tail;
    if (lane < 63) piece_add(total, S1);
    if (lane == 63 || tc == pass63) {
      s_key[2 * wv + 1] = tc;  // one such cell per wave
      s_val[2 * wv + 1] = total;
    } else {
      piece_store(cell_box, cell_flags, tc, total);
    }
  }
  if (lane == 0 && link_in) {  // a cell that started before the wave
    s_key[2 * wv] = hc;
    s_val[2 * wv] = S;
  }
  // the workgroup's slots in point order; equal neighbours are one cell
  __syncthreads();
  if (t == 0) {
    int32_t rc = -1;
    BoxPiece rb = piece_empty();
    for (int q = 0; q < 8; ++q) {
      const int32_t qc = s_key[q];
      if (qc < 0) continue;
      if (qc == rc) {
        piece_add(rb, s_val[q]);
        continue;
      }
      if (rc >= 0) piece_fold(cell_box, cell_flags, rc, rb);
      rc = qc;
      rb = s_val[q];
    }
    if (rc >= 0) piece_fold(cell_box, cell_flags, rc, rb);
  }
}

template <class PP>
__global__ __launch_bounds__(256) void k_bbox(const Params* __restrict__ prm,
                                              const float* __restrict__ xs_a,
                                              const float* __restrict__ ys_a,
                                              const float* __restrict__ xs_b,
                                              const float* __restrict__ ys_b,
                                              const int32_t* __restrict__ point_cell,
                                              const int32_t* __restrict__ cell_start,
                                              uint4* __restrict__ cell_box,
                                              uint32_t* __restrict__ cell_flags,
                                              uint8_t* __restrict__ is_core, int64_t n,
                                              PP pp) {
  bbox_tile(prm, xs_a, ys_a, xs_b, ys_b, point_cell, cell_start, cell_box, cell_flags, is_core,
            n, pp, blockIdx.x);
}

// k_bbox and k_ngh read only what k_cells wrote and write disjoint arrays, so
// inputs of up to kFuseBoxNghMax points run them as one launch: its
// workgroups alternate between a box tile and a block of window searches
// while both have work (the streaming box pass and the latency-bound searches
// share the CUs, and one kernel boundary goes), then take the rest of the
// larger one.  The roles branch on the workgroup id, uniformly for the whole
// workgroup.  On one MI355X the single launch was 1-5% faster end to end for
// 0.2M-3M points and 0.6-2% slower for 10M points and a 135M-point batch.
constexpr int64_t kFuseBoxNghMax = int64_t{1} << 22;
template <bool kSeg, class PP>
__global__ __launch_bounds__(256) void k_bbox_ngh(
    Params* __restrict__ prm, const float* __restrict__ xs_a, const float* __restrict__ ys_a,
    const float* __restrict__ xs_b, const float* __restrict__ ys_b,
    const int32_t* __restrict__ point_cell, const int32_t* __restrict__ cell_start,
    uint4* __restrict__ cell_box, uint32_t* __restrict__ cell_flags,
    uint8_t* __restrict__ is_core, int64_t n, PP pp, uint32_t box_tiles,
    const uint64_t* __restrict__ cell_key, const int2* __restrict__ col_range, int64_t max_cols,
    int32_t* __restrict__ ngh, int32_t* __restrict__ cell_ref, int32_t* __restrict__ uf,
    const int32_t* __restrict__ cell_cloud, const CloudInfo* __restrict__ clouds,
    float* __restrict__ spare_xa, float* __restrict__ spare_xb, uint32_t ngh_blocks) {
  const uint32_t b = blockIdx.x, m = min(box_tiles, ngh_blocks);
  bool box;
  uint32_t i;
  if (b < 2 * m) {
    box = (b & 1u) == 0;
    i = b >> 1;
  } else {
    box = box_tiles > ngh_blocks;
    i = b - m;
  }
  if (box)
    bbox_tile(prm, xs_a, ys_a, xs_b, ys_b, point_cell, cell_start, cell_box, cell_flags, is_core,
              n, pp, i);
  else
    ngh_cells<kSeg>(prm, cell_key, col_range, max_cols, ngh, cell_ref, uf, cell_cloud, clouds,
                    spare_xa, spare_xb, i, ngh_blocks);
}

// Cells up to this many points are scanned by a single lane; larger cells are
// handed to the whole wave.
constexpr int32_t kLaneScan = 64;

// Counts the points of [cs, ce) within eps of (x, y) for one lane, four
// candidates per step, until `total` reaches ms.
__device__ __forceinline__ int32_t lane_count_range(const float* __restrict__ xs,
                                                    const float* __restrict__ ys, int32_t cs,
                                                    int32_t ce, float x, float y, float eps_sq,
                                                    int32_t ms, int32_t total) {
  int32_t j = cs;
  for (; j + 4 <= ce && total < ms; j += 4) {
    float px[4], py[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      px[u] = xs[j + u];
      py[u] = ys[j + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float ex = px[u] - x, ey = py[u] - y;
      total += (ex * ex + ey * ey <= eps_sq) ? 1 : 0;
    }
  }
  for (; j < ce && total < ms; ++j) {
    const float ex = xs[j] - x, ey = ys[j] - y;
    total += (ex * ex + ey * ey <= eps_sq) ? 1 : 0;
  }
  return total;
}

// Smallest d2 <= eps_sq from (x, y) to a core point of [cs, ce), as float
// bits (0xFFFFFFFF if none), for one lane, four candidates per step.
__device__ __forceinline__ uint32_t lane_min_core_d2(const float* __restrict__ xs,
                                                     const float* __restrict__ ys,
                                                     const uint8_t* __restrict__ is_core,
                                                     int32_t cs, int32_t ce, float x, float y,
                                                     float eps_sq) {
  uint32_t m = 0xFFFFFFFFu;
  int32_t j = cs;
  for (; j + 4 <= ce; j += 4) {
    float px[4], py[4];
    uint8_t pc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      px[u] = xs[j + u];
      py[u] = ys[j + u];
      pc[u] = is_core[j + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float ex = px[u] - x, ey = py[u] - y;
      const float d2 = ex * ex + ey * ey;
      if (pc[u] && d2 <= eps_sq) m = min(m, __float_as_uint(d2));
    }
  }
  for (; j < ce; ++j) {
    const float ex = xs[j] - x, ey = ys[j] - y;
    const float d2 = ex * ex + ey * ey;
    if (is_core[j] && d2 <= eps_sq) m = min(m, __float_as_uint(d2));
  }
  return m;
}
// A first try for a cell pair by one lane: the first kSampleSide points of
// each cell against each other (cores only).  A hit is a core pair within eps,
// which decides the pair; otherwise the whole wave tests it.  Adjacent cells
// inside a dense region nearly always connect at the first try.  Small pairs
// take the same route: one lane scanning a whole pair (up to 1,024 point
// pairs) holds its wave for its slowest lane, and on one MI355X that rule was
// slower on every workload measured (the 2,048-cloud batch by 5% end to end).
constexpr int kSampleSide = 4;

__device__ __forceinline__ bool core_pair_sample(const float* __restrict__ xs,
                                                 const float* __restrict__ ys,
                                                 const uint8_t* __restrict__ is_core, int32_t sa,
                                                 int32_t ea, int32_t sb, int32_t eb,
                                                 float eps_sq) {
  float ax[kSampleSide], ay[kSampleSide], bx[kSampleSide], by[kSampleSide];
  bool ac[kSampleSide], bc[kSampleSide];
#pragma unroll
  for (int k = 0; k < kSampleSide; ++k) {
    const int32_t a = sa + k < ea ? sa + k : ea - 1;  // clamped: repeats the last point
    const int32_t b = sb + k < eb ? sb + k : eb - 1;
    ax[k] = xs[a]; ay[k] = ys[a]; ac[k] = is_core[a] != 0;
    bx[k] = xs[b]; by[k] = ys[b]; bc[k] = is_core[b] != 0;
  }
  bool hit = false;
#pragma unroll
  for (int i = 0; i < kSampleSide; ++i) {
#pragma unroll
    for (int k = 0; k < kSampleSide; ++k) {
      const float ex = bx[k] - ax[i], ey = by[k] - ay[i];
      hit |= ac[i] && bc[k] && ex * ex + ey * ey <= eps_sq;
    }
  }
  return hit;
}

// Largest computed squared distance from (x, y) to any point of the box:
// |fl(q - x)| is at most |fl(b - x)| for the box edge b on q's side (monotone
// rounding), so every point of the box has d2 <= this value.  Valid only for
// cells without NaN points (their box ignores them).
__device__ __forceinline__ float far_box_d2(float x, float y, float4 b) {
  const float dx = fmaxf(fabsf(b.x - x), fabsf(b.y - x));
  const float dy = fmaxf(fabsf(b.z - y), fabsf(b.w - y));
  return dx * dx + dy * dy;
}

// Every point of cell nc is within eps of (x, y): box test, finite.  A cell
// holding a NaN point has +inf as its box's x maximum (k_bbox), so it fails.
__device__ __forceinline__ bool cell_inside(float x, float y, float4 b, float eps_sq) {
  const float far = far_box_d2(x, y, b);
  return far <= eps_sq && far < __builtin_inff();
}

// The cells of one window column: at most 5 consecutive cells [j0, j0 + cnt)
// (one per y' in [y'-2, y'+2]).  Their starts, flags and boxes are loaded
// together so that a lane waits once per column rather than once per cell.
struct WinColumn {
  int32_t j0, cnt;
  int32_t start[6];
  uint32_t flags[5];
  float4 box[5];
};

template <bool kFlags = true>
__device__ __forceinline__ void load_column(const int32_t* __restrict__ cell_start,
                                            const uint32_t* __restrict__ cell_flags,
                                            const uint4* __restrict__ cell_box, int32_t j0,
                                            int32_t j1, WinColumn& w) {
  w.j0 = j0;
  w.cnt = j1 > j0 ? j1 - j0 : 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) w.start[k] = (k <= w.cnt && w.cnt > 0) ? cell_start[j0 + k] : 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    if constexpr (kFlags) w.flags[k] = k < w.cnt ? cell_flags[j0 + k] : 0u;
    w.box[k] = k < w.cnt ? load_box(cell_box, j0 + k) : make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Columns of the window nearest first (counting order is free).
__device__ __forceinline__ int near_col(int k) {
  return k == 0 ? 2 : (k == 1 ? 1 : (k == 2 ? 3 : (k == 3 ? 0 : 4)));
}

// Counts, for the point (xi, yi), the points of [p0, p1) with
// fl(fl(ex*ex) + fl(ey*ey)) <= eps_sq, 64 per step, until `total` reaches ms.
__device__ __forceinline__ void count_range(const float* __restrict__ xs,
                                            const float* __restrict__ ys, int32_t p0, int32_t p1,
                                            float xi, float yi, float eps_sq, int32_t ms,
                                            int32_t& total) {
  const int lane = threadIdx.x & 63;
  for (int32_t base = p0; base < p1 && total < ms; base += 64) {
    const int32_t j = base + lane;
    bool hit = false;
    if (j < p1) {
      const float ex = xs[j] - xi, ey = ys[j] - yi;
      hit = ex * ex + ey * ey <= eps_sq;
    }
    total += static_cast<int32_t>(__popcll(__ballot(hit)));
  }
}

// Pass-2 count for one point by one lane (dbscan_grid.hpp:284-307), one
// window column [j0, j1) at a time: window cells whose box is farther than
// eps add nothing, cells whose box lies within eps add all their points,
// small cells are scanned (the own cell c was counted wholesale by the
// caller).  Sets `deferred` when a big cell would need a scan.
__device__ __forceinline__ int32_t lane_column_count(
    const float* __restrict__ xs, const float* __restrict__ ys, float x, float y, int32_t c,
    int32_t j0, int32_t j1, const int32_t* __restrict__ cell_start,
    const uint32_t* __restrict__ cell_flags, const uint4* __restrict__ cell_box,
    int32_t min_samples, float eps_sq, int32_t total, bool& deferred) {
  WinColumn w;
  load_column<false>(cell_start, cell_flags, cell_box, j0, j1, w);  // flags unused
  // the column's cells nearest box first (the count stops at min_samples):
  // cells within eps wholesale, then the partial ones by box distance
  // (eps_sq may be +inf: the cells still to scan are a bit mask, not a
  // distance sentinel)
  float d[5];
  uint32_t todo = 0;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    d[q] = 0.f;
    if (q >= w.cnt || w.j0 + q == c) continue;
    d[q] = point_box_d2(x, y, w.box[q]);
    if (d[q] > eps_sq) continue;
    if (cell_inside(x, y, w.box[q], eps_sq))
      total += w.start[q + 1] - w.start[q];
    else
      todo |= 1u << q;
  }
#pragma unroll
  for (int it = 0; it < 5; ++it) {
    if (!todo || total >= min_samples) break;
    int qm = -1;
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (((todo >> q) & 1u) && (qm < 0 || d[q] < d[qm])) qm = q;
    todo &= ~(1u << qm);
    int32_t cs = 0, ce = 0;
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (q == qm) {
        cs = w.start[q];
        ce = w.start[q + 1];
      }
    if (ce - cs <= kLaneScan)
      total = lane_count_range(xs, ys, cs, ce, x, y, eps_sq, min_samples, total);
    else
      deferred = true;
  }
  return total;
}

// The same count by the whole wave for one point (xi, yi) of cell ci: small
// cells are scanned in runs, big cells after their box tests, 64 candidates
// per step, until min_samples.
__device__ int32_t wave_core_count(const float* __restrict__ xs, const float* __restrict__ ys,
                                   float xi, float yi, int32_t ci,
                                   const int32_t* __restrict__ row,
                                   const int32_t* __restrict__ cell_start,
                                   const uint32_t* __restrict__ cell_flags,
                                   const uint4* __restrict__ cell_box, int32_t min_samples,
                                   float eps_sq, int32_t total) {
  for (int k = 0; k < 5 && total < min_samples; ++k) {
    const int col = near_col(k);
    const int32_t j0 = row[2 * col], j1 = row[2 * col + 1];
    if (j0 >= j1) continue;
    int32_t run = cell_start[j0];
    for (int32_t nc = j0; nc < j1 && total < min_samples; ++nc) {
      const int32_t cs = cell_start[nc], ce = cell_start[nc + 1];
      const bool own = nc == ci;
      if (!own && ce - cs <= kLaneScan) continue;
      count_range(xs, ys, run, cs, xi, yi, eps_sq, min_samples, total);
      run = ce;
      if (own) continue;
      const float4 b = load_box(cell_box, nc);
      if (point_box_d2(xi, yi, b) > eps_sq) continue;
      if (cell_inside(xi, yi, b, eps_sq)) {
        total += ce - cs;
        continue;
      }
      count_range(xs, ys, cs, ce, xi, yi, eps_sq, min_samples, total);
    }
    count_range(xs, ys, run, cell_start[j1], xi, yi, eps_sq, min_samples, total);
  }
  return total;
}


// Cells a wave takes per round in the per-cell passes (one lane each): 64
// when there are enough cells to keep the GPU's wave slots (kWaveSlots,
// about 256 CUs x 32) busy with full rounds, else a smaller power of two,
// down to 8.  The pass-2 points of a round are shared out over all 64 lanes
// whatever the count, so a small input's points spread over many waves
// instead of queueing in a few; past that size, fewer cells per wave only
// leave lanes idle.
constexpr int64_t kWaveSlots = 8192;
__device__ __forceinline__ int64_t cells_per_wave(int64_t C) {
  const int64_t per = C / kWaveSlots;
  return per > 32 ? 64 : per > 16 ? 32 : per > 8 ? 16 : 8;
}

// Core marking pass 2 (dbscan_grid.hpp:281-307) for the cells with fewer than
// min_samples points (pass 1 was done by k_bbox).  A wave takes up to 64
// cells (one lane each), prunes the cells whose whole window holds fewer than
// min_samples points, and counts for the remaining cells' points from a work
// list, one window column per lane and step; the counts stop at min_samples,
// which cannot change the outcome.  Points whose count needs a scan of a big
// cell, and cells of more than 64 points (possible when min_samples > 64),
// are counted by the whole wave.  Sets the cell's core bits, once per cell.
// Launch bounds hold it to 72 VGPRs (7 waves per SIMD; 74 uncapped gave 6:
// 6% faster on skewed10m).
template <class PP>
__global__ __launch_bounds__(256, 7) void k_core_sparse(
    const Params* __restrict__ prm, const float* __restrict__ xs_a,
    const float* __restrict__ ys_a, const float* __restrict__ xs_b,
    const float* __restrict__ ys_b, const int32_t* __restrict__ cell_start,
    const int32_t* __restrict__ ngh, const uint4* __restrict__ cell_box,
    uint8_t* __restrict__ is_core, uint32_t* __restrict__ cell_flags, PP pp) {
  const int64_t C = prm->num_cells;
  const bool in_a = (prm->npasses & 1) != 0;
  const float* __restrict__ xs = in_a ? xs_a : xs_b;
  const float* __restrict__ ys = in_a ? ys_a : ys_b;
  const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
  const int64_t wave = (static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x) >> 6;
  const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * 256) >> 6;
  const int64_t K = cells_per_wave(C);
  // per wave: the window row of the cell each lane owns (column-major, so
  // that the lanes' words fall in distinct banks; read by the lanes counting
  // its points) and the core / non-core bits of that cell, gathered here and
  // written once per cell
  __shared__ int32_t s_row[4][10][64];
  __shared__ uint32_t s_bits[4][64];
  for (int64_t base = wave * K; base < C; base += nwaves * K) {
    const int32_t c = static_cast<int32_t>(base + lane);
    const bool valid = lane < K && base + lane < C;
    int32_t s = 0, e = 0;
    CellParams cp{0.f, 0.f, 0};
    if (valid) {
      s = cell_start[c];
      e = cell_start[c + 1];
      cp = pp.at(c);
    }
    s_bits[wib][lane] = 0u;
    const int32_t min_samples = cp.ms;
    const float eps_sq = cp.eps_sq;
    const bool sparse = valid && e - s < min_samples;
    const bool whole_wave = sparse && e - s > 64;
    uint32_t bits = 0;
    bool lonely = false;  // fewer than min_samples points in the whole window
    if (sparse && !whole_wave) {
      const int32_t* row = ngh + static_cast<int64_t>(c) * 10;
      // A point's count never exceeds the points of its 25 window cells, the
      // five column runs [j0, j1): if they hold fewer than min_samples, no
      // point of the cell is core.  (Exact, and one round of loads.)
      int32_t j[10];
#pragma unroll
      for (int k = 0; k < 10; ++k) j[k] = row[k];
      int32_t wt = 0;
#pragma unroll
      for (int k = 0; k < 5; ++k)
        wt += j[2 * k + 1] > j[2 * k] ? cell_start[j[2 * k + 1]] - cell_start[j[2 * k]] : 0;
      lonely = wt < min_samples;
      // the row, for the lanes that will count this cell's points
#pragma unroll
      for (int k = 0; k < 10; ++k) s_row[wib][k][lane] = j[k];
      if (lonely) {
        for (int32_t p = s; p < e; ++p) is_core[p] = 0;
        bits = kCellHasNonCore;
      }
    }
    __builtin_amdgcn_wave_barrier();  // the owners' rows are in LDS
    // The points of the cells still undecided form the wave's work list:
    // item q is point q - excl of the cell of the lane whose [excl, incl)
    // holds q (a binary search over the lanes' prefix sums).  A lane takes an
    // item and counts for it one window column per step, nearest column
    // first.  A lane whose point is decided (min_samples reached, or all five
    // columns seen) takes the next item in the same step, so every step
    // costs one column's loads for the lanes at work, instead of a round of
    // 64 points waiting for the lane whose point needs the most columns.
    const uint32_t cnt = (sparse && !whole_wave && !lonely) ? static_cast<uint32_t>(e - s) : 0u;
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
    uint32_t next = 0;  // the first item not yet taken (wave-uniform)
    bool need = true, active = false, deferred = false;
    int32_t p = 0, oc = 0, oms = 0, tot = 0, k = 0, ol = 0;
    float px = 0.f, py = 0.f, oeps = 0.f;
    if constexpr (PP::kUniform) {  // kernel arguments, no per-lane copies
      oms = pp.at(0).ms;
      oeps = pp.at(0).eps_sq;
    }
    while (next < total || __ballot(active)) {
      const uint64_t nm = __ballot(need);
      if (nm && next >= total) {
        need = false;  // nothing left to take: these lanes are done
      } else if (nm) {
        const uint32_t q = next + static_cast<uint32_t>(
                                      __popcll(nm & ((1ull << lane) - 1ull)));
        next += static_cast<uint32_t>(__popcll(nm));
        int o = 0;  // owner lane of item q: the number of lanes with incl <= q
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
          const uint32_t v = static_cast<uint32_t>(__shfl(static_cast<int>(incl), o + step - 1, 64));
          if (v <= q) o += step;
        }
        o = min(o, 63);
        const int32_t c_o = __shfl(c, o, 64), s_o = __shfl(s, o, 64), e_o = __shfl(e, o, 64);
        const uint32_t x_o = static_cast<uint32_t>(__shfl(static_cast<int>(incl - cnt), o, 64));
        int32_t ms_o = 0;
        float eps_o = 0.f;
        if constexpr (!PP::kUniform) {
          ms_o = __shfl(min_samples, o, 64);
          eps_o = __shfl(eps_sq, o, 64);
        }
        if (need) {
          need = false;
          active = q < total;
          if (active) {
            ol = o;
            oc = c_o;
            if constexpr (!PP::kUniform) {
              oms = ms_o;
              oeps = eps_o;
            }
            p = s_o + static_cast<int32_t>(q - x_o);
            tot = e_o - s_o;  // the own cell, wholesale
            k = 0;
            deferred = false;
            px = xs[p];  // used only after the first column's loads are issued
            py = ys[p];
          }
        }
      }
      bool finished = false;
      if (active) {
        const int col = near_col(k);
        const int32_t j0 = s_row[wib][2 * col][ol], j1 = s_row[wib][2 * col + 1][ol];
        tot = lane_column_count(xs, ys, px, py, oc, j0, j1, cell_start, cell_flags, cell_box, oms,
                                oeps, tot, deferred);
        ++k;
        if (tot >= oms || k == 5) {
          finished = true;
          active = false;
          need = true;
          if (!deferred || tot >= oms) {
            const bool core = tot >= oms;
            is_core[p] = core ? 1 : 0;
            atomicOr(&s_bits[wib][ol], core ? kCellHasCore : kCellHasNonCore);
            deferred = false;
          }
        }
      }
      // points whose count needs a scan of a big cell: the whole wave, one by one
      uint64_t dm = __ballot(finished && deferred);
      while (dm) {
        const int src = lowest_lane(dm);
        dm &= dm - 1ull;
        const int32_t pc = bcast(oc, src), pp_ = bcast(p, src);
        const int32_t pms = bcast(oms, src);
        const float peps = bcastf(oeps, src);
        const int32_t* row = ngh + static_cast<int64_t>(pc) * 10;
        const int32_t wt = wave_core_count(xs, ys, xs[pp_], ys[pp_], pc, row, cell_start,
                                           cell_flags, cell_box, pms, peps,
                                           cell_start[pc + 1] - cell_start[pc]);
        if (lane == src) {
          const bool core = wt >= pms;
          is_core[pp_] = core ? 1 : 0;
          atomicOr(&s_bits[wib][ol], core ? kCellHasCore : kCellHasNonCore);
        }
      }
    }
    // the wave's own LDS writes, in order: every point's bits are in
    __builtin_amdgcn_wave_barrier();
    if (cnt) {
      const uint32_t b = s_bits[wib][lane];
      if (b) atomicOr(cell_flags + c, b);
    }
    // cells of more than 64 points (min_samples > 64): the whole wave
    uint64_t todo = __ballot(whole_wave);
    while (todo) {
      const int src = lowest_lane(todo);
      todo &= todo - 1ull;
      const int32_t cc = bcast(c, src), cs = bcast(s, src), ce = bcast(e, src);
      const int32_t wms = bcast(min_samples, src);
      const float weps = bcastf(eps_sq, src);
      const int32_t* row = ngh + static_cast<int64_t>(cc) * 10;
      for (int32_t p = cs; p < ce; ++p) {
        const int32_t tot = wave_core_count(xs, ys, xs[p], ys[p], cc, row, cell_start,
                                            cell_flags, cell_box, wms, weps, ce - cs);
        if (lane == src) {
          const bool core = tot >= wms;
          is_core[p] = core ? 1 : 0;
          bits |= core ? kCellHasCore : kCellHasNonCore;
        }
      }
    }
    if (lonely || whole_wave) cell_flags[c] |= bits;  // the lane owns these cells' bits
  }
}

// True if some core point of [sa, ea) and some core point of [sb, eb) are
// within eps (dbscan_grid.hpp:370-384), for the whole wave: points are first
// filtered by their distance to the other cell's box, the surviving pairs are
// tested 64 at a time, and the search stops at the first hit.
__device__ bool core_pair_within(const float* __restrict__ xs, const float* __restrict__ ys,
                                 const uint8_t* __restrict__ is_core, int32_t sa, int32_t ea,
                                 float4 box_a, int32_t sb, int32_t eb, float4 box_b,
                                 float eps_sq) {
  const int lane = threadIdx.x & 63;
  for (int32_t ba = sa; ba < ea; ba += 64) {
    const int32_t a = ba + lane;
    float xa = 0.f, ya = 0.f;
    bool va = false;
    if (a < ea && is_core[a]) {
      xa = xs[a];
      ya = ys[a];
      va = point_box_d2(xa, ya, box_b) <= eps_sq;
    }
    const uint64_t ma = __ballot(va);
    if (!ma) continue;
    for (int32_t bb = sb; bb < eb; bb += 64) {
      const int32_t b = bb + lane;
      float xb = 0.f, yb = 0.f;
      bool vb = false;
      if (b < eb && is_core[b]) {
        xb = xs[b];
        yb = ys[b];
        vb = point_box_d2(xb, yb, box_a) <= eps_sq;
      }
      if (!__ballot(vb)) continue;
      uint64_t m = ma;
      while (m) {
        const int src = lowest_lane(m);
        m &= m - 1ull;
        const float xi = bcastf(xa, src), yi = bcastf(ya, src);
        const float ex = xb - xi, ey = yb - yi;
        if (__ballot(vb && ex * ex + ey * ey <= eps_sq)) return true;
      }
    }
  }
  return false;
}

// Cell-level union over eps-close core pairs (dbscan_grid.hpp:351-387).  One
// lane per (cell c, forward neighbour slot): each unordered window pair is
// visited once, from its smaller internal index (the rest of c's column above
// c, and the two columns to the right; at most 12 slots).  Pairs with a core
// on both sides and boxes within eps are sampled by the lane
// (core_pair_sample) and, when that does not decide them, tested by the whole
// wave.
// Two launches split the pairs: kPhase 0 takes the 3x3 neighbours (dx <= 1,
// |dy| <= 1; only slots 0-6 can hold one), which join almost every dense
// region, and only records each hit as bit `slot` of the cell's mask (the
// unions follow in k_union_local, mostly in LDS); after a flatten, kPhase 1
// takes the rest, first skipping pairs whose cells already share a root, and
// unites in uf directly.  The union is order-independent, so the split cannot
// change a label.
constexpr int kUnionSlots = 12;
// Cells per workgroup of k_union_local; its ranges start at multiples of it.
constexpr int kUnionRange = 1024;

// Forward neighbour `slot` of cell c, with its column offset, or -1.  Slots
// 0-1: the next two cells of c's column (ra = the end of c's window run in
// its own column, window-row word 5); 2-6: the cells of the column to the
// right, 7-11: the column after that ([ra, rb) = that column's window run,
// words 6-7 or 8-9).
__device__ __forceinline__ int32_t forward_neighbour(int32_t c, int slot, int32_t ra, int32_t rb,
                                                     int& dx) {
  if (slot < 2) {
    dx = 0;
    return c + 1 + slot < ra ? c + 1 + slot : -1;
  }
  dx = slot < 7 ? 1 : 2;
  const int32_t k = ra + (slot - (slot < 7 ? 2 : 7));
  return k < rb ? k : -1;
}

// Launch bounds hold it to 64 VGPRs (8 waves per SIMD; 66 uncapped gave 7):
// the pair tests wait on loads, and the extra wave was 0.2-1.3% faster end to
// end on one MI355X.
template <int kPhase, class PP>
__global__ __launch_bounds__(256, 8) void k_union(
    Params* __restrict__ prm, const float* __restrict__ xs_a, const float* __restrict__ ys_a,
    const float* __restrict__ xs_b, const float* __restrict__ ys_b,
    const int32_t* __restrict__ cell_start, const int32_t* __restrict__ ngh,
    const int32_t* __restrict__ cell_ref, const uint32_t* __restrict__ cell_flags,
    const uint4* __restrict__ cell_box, const uint64_t* __restrict__ cell_key,
    const uint8_t* __restrict__ is_core, int32_t* __restrict__ uf, PP pp,
    float* __restrict__ spare_xa, float* __restrict__ spare_xb) {
  constexpr int S = kPhase == 0 ? 7 : kUnionSlots;  // slots scanned by this phase
  uint32_t* __restrict__ hits = spare_words(prm, spare_xa, spare_xb);
  const int64_t C = prm->num_cells;
  const bool in_a = (prm->npasses & 1) != 0;
  const float* __restrict__ xs = in_a ? xs_a : xs_b;
  const float* __restrict__ ys = in_a ? ys_a : ys_b;
  const int lane = threadIdx.x & 63;
  const int64_t wave = (static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x) >> 6;
  const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * 256) >> 6;
  const int64_t total = C * S;
  for (int64_t base = wave * 64; base < total; base += nwaves * 64) {
    const int64_t t = base + lane;
    int32_t c = -1, nc = -1;
    int slot = 0;
    // Three rounds of loads: the lane's own cell with its window run for
    // `slot`, then the neighbour, then the point samples.
    int32_t sa = 0, ea = 0, sb = 0, eb = 0;
    int32_t root_c = -1, ref_n = -1;  // kPhase 1: c's root, nc's reference index
    float4 box_c = make_float4(0.f, 0.f, 0.f, 0.f), box_n = box_c;
    if (t < total) {
      c = static_cast<int32_t>(t / S);
      slot = static_cast<int>(t - static_cast<int64_t>(c) * S);
      const int32_t* row = ngh + static_cast<int64_t>(c) * 10;
      const int ia = slot < 2 ? 5 : (slot < 7 ? 6 : 8);
      const int32_t ra = row[ia], rb = row[slot < 2 ? 5 : ia + 1];
      const uint32_t fl_c = cell_flags[c];
      const uint32_t ky_c = static_cast<uint32_t>(cell_key[c]);
      sa = cell_start[c];
      ea = cell_start[c + 1];
      box_c = load_box(cell_box, c);
      int32_t ref_c = 0;
      if constexpr (kPhase == 1) ref_c = cell_ref[c];
      if (fl_c & kCellHasCore) {
        int dx;
        nc = forward_neighbour(c, slot, ra, rb, dx);
        if (nc >= 0) {
          if constexpr (kPhase == 1) {
            root_c = uf[ref_c];  // with the neighbour's loads
            ref_n = cell_ref[nc];
          }
          const uint32_t ky_n = static_cast<uint32_t>(cell_key[nc]);
          const uint32_t fl_n = cell_flags[nc];
          box_n = load_box(cell_box, nc);
          sb = cell_start[nc];
          eb = cell_start[nc + 1];
          const int64_t dy = static_cast<int64_t>(ky_n) - static_cast<int64_t>(ky_c);
          if ((dx <= 1 && dy >= -1 && dy <= 1) != (kPhase == 0) || !(fl_n & kCellHasCore))
            nc = -1;
        }
      }
    }
    // after the first phase's flatten, uf holds roots: equal roots need no test
    if (kPhase == 1 && nc >= 0 && root_c == uf[ref_n]) nc = -1;
    bool deferred = false;
    float eps_sq = 0.f;  // c and nc are cells of one cloud
    if (nc >= 0) {
      eps_sq = pp.at(c).eps_sq;
      if (box_box_d2(box_c, box_n) <= eps_sq) {
        const bool hit = core_pair_sample(xs, ys, is_core, sa, ea, sb, eb, eps_sq);
        deferred = !hit;
        if (hit) {
          if constexpr (kPhase == 0)
            atomicOr(hits + c, 1u << slot);
          else
            uf_unite(uf, cell_ref[c], cell_ref[nc], &prm->error);
        }
      }
    }
    uint64_t todo = __ballot(deferred);
    while (todo) {
      const int src = lowest_lane(todo);
      todo &= todo - 1ull;
      const int32_t cw = bcast(c, src), nw = bcast(nc, src), sw = bcast(slot, src);
      const float weps = bcastf(eps_sq, src);
      const int32_t rc = cell_ref[cw], rn = cell_ref[nw];
      if (kPhase == 1 && uf_find_ro(uf, rc, &prm->error) == uf_find_ro(uf, rn, &prm->error))
        continue;
      const bool hit = core_pair_within(xs, ys, is_core, cell_start[cw], cell_start[cw + 1],
                                        load_box(cell_box, cw), cell_start[nw],
                                        cell_start[nw + 1], load_box(cell_box, nw), weps);
      if (hit && lane == 0) {
        if constexpr (kPhase == 0)
          atomicOr(hits + cw, 1u << sw);
        else
          uf_unite(uf, rc, rn, &prm->error);
      }
    }
  }
}

// The unions of the 3x3 hits (masks from k_union<0>).  A workgroup takes a
// range of kUnionRange consecutive cells (a few whole columns of a dense
// region, so nearly every 3x3 neighbour is in the same range): hits between
// two of its cells are unions in LDS (parents in local indices, hooked by
// reference index, larger onto smaller).  The range's forest then goes to uf,
// each linked cell onto its local root, by CAS (a cell that another
// workgroup's cross-range union hooked meanwhile takes a union in uf
// instead); only then are the hits reaching past the range united in uf, so
// that their finds start from written forests and repeated hits between two
// components end at the first step.  Every step links a root onto a smaller
// reference index, so any interleaving gives the same components.  The pair
// tests stay in k_union<0>, spread evenly over the grid; this pass only reads
// masks and window rows.
constexpr int kUnionLocalThreads = 512;
constexpr int kCrossCap = 2048;

__device__ __forceinline__ int32_t lds_find(int32_t* s_par, int32_t x) {
  int32_t p = __hip_atomic_load(s_par + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  while (p != x) {
    const int32_t gp =
        __hip_atomic_load(s_par + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (gp != p)
      __hip_atomic_store(s_par + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    x = p;
    p = gp;
  }
  return x;
}

__device__ void lds_unite(int32_t* s_par, const int32_t* s_ref, int32_t a, int32_t b,
                          int32_t* error) {
  for (uint32_t it = 0; it < kSpinLimit; ++it) {
    a = lds_find(s_par, a);
    b = lds_find(s_par, b);
    if (a == b) return;
    if (s_ref[a] < s_ref[b]) {
      const int32_t tmp = a;
      a = b;
      b = tmp;
    }
    if (atomicCAS(s_par + a, a, b) == a) return;
  }
  atomicOr(error, static_cast<int32_t>(kErrUfLoop));
}

__global__ __launch_bounds__(kUnionLocalThreads) void k_union_local(
    Params* __restrict__ prm, const int32_t* __restrict__ ngh,
    const int32_t* __restrict__ cell_ref, int32_t* __restrict__ uf,
    float* __restrict__ spare_xa, float* __restrict__ spare_xb) {
  __shared__ int32_t s_par[kUnionRange];
  __shared__ int32_t s_ref[kUnionRange];
  __shared__ int2 s_cross[kCrossCap];  // (local cell, neighbour) of hits past the range
  __shared__ int32_t s_ncross;
  const uint32_t* __restrict__ hits = spare_words(prm, spare_xa, spare_xb);
  const int64_t C = prm->num_cells;
  const int t = threadIdx.x;
  for (int64_t c0 = static_cast<int64_t>(blockIdx.x) * kUnionRange; c0 < C;
       c0 += static_cast<int64_t>(gridDim.x) * kUnionRange) {
    const int32_t nloc = static_cast<int32_t>(min<int64_t>(kUnionRange, C - c0));
    const int64_t c1 = c0 + nloc;
    for (int i = t; i < nloc; i += kUnionLocalThreads) {
      s_par[i] = i;
      s_ref[i] = cell_ref[c0 + i];
    }
    if (t == 0) s_ncross = 0;
    __syncthreads();
    for (int i = t; i < nloc; i += kUnionLocalThreads) {
      const int32_t c = static_cast<int32_t>(c0) + i;
      uint32_t m = hits[c];
      // the 3x3 slots (0-6) need only the window row's words 5-7: loaded
      // once for the cell, not once per hit
      int32_t r5 = 0, r6 = 0, r7 = 0;
      if (m) {
        const int32_t* row = ngh + static_cast<int64_t>(c) * 10;
        r5 = row[5];
        r6 = row[6];
        r7 = row[7];
      }
      while (m) {
        const int slot = __builtin_ctz(m);
        m &= m - 1u;
        int dx;
        const int32_t nc = slot < 2 ? forward_neighbour(c, slot, r5, r5, dx)
                                    : forward_neighbour(c, slot, r6, r7, dx);
        if (nc < 0 || slot >= 7) {  // a hit is only recorded for an existing neighbour
          atomicOr(&prm->error, static_cast<int32_t>(kErrRange));
        } else if (nc < c1) {
          lds_unite(s_par, s_ref, i, nc - static_cast<int32_t>(c0), &prm->error);
        } else {
          const int32_t q = atomicAdd(&s_ncross, 1);
          if (q < kCrossCap)
            s_cross[q] = make_int2(i, nc);
          else  // list full: united now (correct in any order, only slower)
            uf_unite(uf, s_ref[i], cell_ref[nc], &prm->error);
        }
      }
    }
    __syncthreads();
    // the range's forest to uf: each linked cell onto its local root
    for (int i = t; i < nloc; i += kUnionLocalThreads) {
      const int32_t r = lds_find(s_par, i);
      if (r != i) {
        const int32_t gx = s_ref[i], gy = s_ref[r];
        if (atomicCAS(uf + gx, gx, gy) != gx) uf_unite(uf, gx, gy, &prm->error);
      }
    }
    __syncthreads();
    const int32_t ncross = min(s_ncross, kCrossCap);
    for (int q = t; q < ncross; q += kUnionLocalThreads) {
      const int2 h = s_cross[q];
      uf_unite(uf, s_ref[h.x], cell_ref[h.y], &prm->error);
    }
    __syncthreads();
  }
}

// uf[r] = root by a read-only walk (each thread writes only its own slot; the
// reference's determinism fix, csrc/cuda/dbscan_cuda.cu:288-299,372-378), and
// the root flag of each reference index.
__global__ __launch_bounds__(256) void k_flatten(Params* __restrict__ prm,
                                                 const int32_t* __restrict__ cell_ref,
                                                 const uint32_t* __restrict__ cell_flags,
                                                 int32_t* __restrict__ uf,
                                                 uint32_t* __restrict__ core_ref) {
  const int32_t C = prm->num_cells;
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; c < C;
       c += static_cast<int64_t>(gridDim.x) * 256) {
    const int32_t r = cell_ref[c];
    const int32_t root = uf_find_ro(uf, r, &prm->error);
    core_ref[r] = cell_flags[c] & 1u;  // the cell's core bit, by reference index
    st_relaxed(uf + r, root);
  }
}

// Exclusive scan of the root flags over reference indices, in place: the
// rank of a root is its cluster id (dbscan_grid.hpp:392-398).  On entry
// flag_rank holds each reference index's core bit (k_flatten); a root is a
// core cell r with uf[r] == r (the unions after the flatten leave uf
// unflattened; its roots are still its fixed points).
__global__ __launch_bounds__(kScanThreads) void k_rank(Params* __restrict__ prm,
                                                       uint64_t* __restrict__ status,
                                                       uint32_t* __restrict__ flag_rank,
                                                       const int32_t* __restrict__ uf) {
  __shared__ uint32_t s_scan[kScanThreads / 64];
  __shared__ uint64_t s_prefix;
  const int64_t C = prm->num_cells;
  const int t = threadIdx.x;
  // blocks past the cell count leave before taking a tile (the grid is sized
  // for n points; tiles are only needed for C cells)
  if (static_cast<int64_t>(blockIdx.x) * kScanTile >= C) return;
  const uint32_t tile = blockIdx.x;  // see kTileOrder
  const int64_t base = static_cast<int64_t>(tile) * kScanTile;
  const int64_t e0 = base + static_cast<int64_t>(t) * kScanItems;
  // uf and the core flags come from earlier kernels: plain loads, both
  // issued together (a root is a core cell that is its own parent)
  uint32_t f[kScanItems];
  int32_t u[kScanItems];
  uint32_t cnt = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int64_t i = e0 + j < C ? e0 + j : C - 1;
    u[j] = uf[i];
    f[j] = flag_rank[i];
  }
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    f[j] = (e0 + j < C && u[j] == static_cast<int32_t>(e0 + j)) ? f[j] : 0u;
    cnt += f[j];
  }
  uint32_t ttot;
  const uint32_t texcl = block_excl_scan<kScanThreads>(cnt, ttot, s_scan);
  if (t < 64) {
    const uint64_t pre = lookback_wave(status, tile, ttot, &prm->error);
    if (t == 0) s_prefix = pre;
  }
  __syncthreads();
  uint32_t run = static_cast<uint32_t>(s_prefix) + texcl;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (e0 + j < C) flag_rank[e0 + j] = run;
    run += f[j];
    if (e0 + j == C - 1) prm->num_roots = static_cast<int32_t>(run);
  }
}

// kSeg (batches): ids count from 0 in every cloud, so the number of roots
// before the cloud's first reference index is subtracted.
template <bool kSeg>
__global__ __launch_bounds__(256) void k_cluster(Params* __restrict__ prm,
                                                 const int32_t* __restrict__ cell_ref,
                                                 const uint32_t* __restrict__ cell_flags,
                                                 int32_t* __restrict__ uf,
                                                 const uint32_t* __restrict__ rank,
                                                 int32_t* __restrict__ cell_cluster,
                                                 const int32_t* __restrict__ cell_cloud,
                                                 const CloudInfo* __restrict__ clouds) {
  const int32_t C = prm->num_cells;
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; c < C;
       c += static_cast<int64_t>(gridDim.x) * 256) {
    int32_t id = kNoise;
    // the root by a read-only walk, written back (uf ends flattened; every
    // value written is a root, so concurrent walks stay on valid ancestors)
    const int32_t r = cell_ref[c];
    const int32_t root = uf_find_ro(uf, r, &prm->error);
    if (root != r) st_relaxed(uf + r, root);
    if (cell_flags[c] & 1u) {
      id = static_cast<int32_t>(rank[root]);
      if constexpr (kSeg) id -= static_cast<int32_t>(rank[clouds[cell_cloud[c]].cell_begin]);
    }
    cell_cluster[c] = id;
  }
}

// Core labels and unpermute (dbscan_grid.hpp:401-412,448-455), in two
// passes so that no label is written alone into a random cache line.
// Buckets of B consecutive input positions (bucket_width(): at most
// kMaxBuckets buckets of at most kUnpermMaxWidth positions, so inputs up to
// 2^24 points; larger ones take k_label_direct): bucket b gathers the points
// whose input position is in [b B, (b+1) B).  B is the smallest multiple of
// 256 (at least 4,096) that needs no more than kMaxBuckets buckets: for 10M
// points 508 buckets of 80 KB of labels, two k_unpermute workgroups per CU,
// one round over the chip.
constexpr int kMaxBuckets = 512;
constexpr int kUnpermMaxWidth = 1 << 15;  // labels k_unpermute holds in LDS at once (128 KB)
constexpr int kUnpermSmallWidth = 20480;  // up to here the 80 KB instantiation
constexpr int kLabelThreads = 1024;
constexpr int kLabelTile = kLabelThreads * 8;  // the smaller tile of the two forms
// With fewer than kPackedRoots roots (so every label + 1 < 2^17) a pair is
// one word: (label + 1) << 15 | position within the bucket (B <= 2^15); else
// two words, input position and label.  The form is read from num_roots on
// the device, in the kernels themselves.
constexpr int32_t kPackedRoots = (1 << 17) - 1;
// One-word pairs go in 12,288-point tiles (48 KB of the 64 KB buffer): at 1M
// to 10M points that was 0.6-1.6% faster end to end than 16,384-point tiles
// on one MI355X (a 10M-point call has 814 tiles instead of 611 for the 512
// workgroup slots, so its last round is less idle).
template <bool kPacked>
struct LabelForm {
  static constexpr int kItems = kPacked ? 12 : 8;
  static constexpr int kTile = kLabelThreads * kItems;
  using Slot = std::conditional_t<kPacked, uint32_t, uint64_t>;  // at most 64 KB of LDS
};
static_assert(LabelForm<true>::kTile * sizeof(uint32_t) <= kLabelTile * sizeof(uint64_t) &&
                  LabelForm<true>::kTile >= kLabelTile,
              "one-word tiles fit the pair buffer and need no more run-table columns");
static_assert(kLabelThreads >= kMaxBuckets, "a thread per bucket");
// A run of more pairs than this marks its bucket for the chunked unpermute.
constexpr uint32_t kLongRun = 256;

// Pass A: every sorted point's label (a core point: its cell's cluster id;
// another point: the border label k_border left in sorted order) and input
// position form a pair.  A tile groups its pairs by bucket in LDS and writes
// them, in that order, to its own slice of the pair arrays (the tile's
// sorted positions), one contiguous write; a table gives where each bucket's
// run starts in the slice: row b, column tile, row nb holding the tile's
// size (the table lives in the radix passes' look-back words, free by now).
// A bucket that gets a run of more than kLongRun pairs is flagged in
// long_runs.  No tile waits for another; the order inside a run is free.
template <bool kPacked>
__device__ __forceinline__ void label_tile(
    Params* __restrict__ prm, const uint32_t* __restrict__ idx_a,
    const uint32_t* __restrict__ idx_b, const int32_t* __restrict__ point_cell,
    const uint8_t* __restrict__ is_core, const int32_t* __restrict__ cell_cluster,
    const uint32_t* __restrict__ spare_ia, const uint32_t* __restrict__ spare_ib,
    uint32_t* __restrict__ spare_xa, uint32_t* __restrict__ spare_xb,
    uint32_t* __restrict__ spare_ya, uint32_t* __restrict__ spare_yb,
    uint32_t* __restrict__ table, uint32_t* __restrict__ long_runs, int64_t n64, uint32_t B,
    uint32_t* s_cnt, typename LabelForm<kPacked>::Slot* s_buf, uint32_t* s_scan) {
  using F = LabelForm<kPacked>;
  constexpr int T = kLabelThreads;
  constexpr int kLabelItems = F::kItems;
  const uint32_t n = static_cast<uint32_t>(n64);
  const bool in_a = (prm->npasses & 1) != 0;
  const uint32_t* __restrict__ idx = in_a ? idx_a : idx_b;
  const int32_t* __restrict__ lab_sorted =
      reinterpret_cast<const int32_t*>(in_a ? spare_ib : spare_ia);
  uint32_t* __restrict__ pair_i = in_a ? spare_xb : spare_xa;
  int32_t* __restrict__ pair_l = reinterpret_cast<int32_t*>(in_a ? spare_yb : spare_ya);
  const int t = threadIdx.x;
  const uint32_t nb = (n + B - 1u) / B;
  const uint32_t tile = blockIdx.x;  // see kTileOrder
  const uint32_t base = tile * F::kTile;
  const uint32_t tile_n = min(static_cast<uint32_t>(F::kTile), n - base);
  for (uint32_t e = t; e < nb; e += T) s_cnt[e] = 0u;
  __syncthreads();
  uint32_t dst[kLabelItems], rk[kLabelItems];
  int32_t lab[kLabelItems];
#pragma unroll
  for (int k = 0; k < kLabelItems; ++k) {
    const uint32_t l = k * T + t;
    dst[k] = n;
    if (l < tile_n) {
      const uint32_t i = base + l;
      dst[k] = idx[i];
      lab[k] = is_core[i] ? cell_cluster[point_cell[i]] : lab_sorted[i];
      if (dst[k] < n) rk[k] = atomicAdd(&s_cnt[dst[k] / B], 1u);
    }
  }
  __syncthreads();
  const bool owner = static_cast<uint32_t>(t) < nb;
  const uint32_t cnt = owner ? s_cnt[t] : 0u;
  uint32_t ttot;
  const uint32_t lstart = block_excl_scan<T>(cnt, ttot, s_scan);
  // run table: row b, column tile = where bucket b's run starts in the tile's
  // slice; row nb = the tile's size (the end of the last run)
  const uint32_t ntiles = (n + F::kTile - 1u) / F::kTile;
  if (owner) {
    table[static_cast<size_t>(t) * ntiles + tile] = lstart;
    s_cnt[t] = lstart;
    if (cnt > kLongRun) long_runs[t] = 1u;
  }
  if (t == 0) table[static_cast<size_t>(nb) * ntiles + tile] = tile_n;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kLabelItems; ++k) {
    if (dst[k] < n) {
      const uint32_t b = dst[k] / B;
      if constexpr (kPacked)
        s_buf[s_cnt[b] + rk[k]] = (static_cast<uint32_t>(lab[k] + 1) << 15) | (dst[k] - b * B);
      else
        s_buf[s_cnt[b] + rk[k]] =
            (static_cast<uint64_t>(static_cast<uint32_t>(lab[k])) << 32) | dst[k];
    } else if (static_cast<uint32_t>(k * T + t) < tile_n) {
      atomicOr(&prm->error, static_cast<int32_t>(kErrRange));
    }
  }
  __syncthreads();
  // the tile's pairs, grouped by bucket, to the tile's own slice
  for (uint32_t j = t; j < tile_n; j += T) {
    const typename F::Slot v = s_buf[j];
    if constexpr (kPacked) {
      pair_i[base + j] = v;
    } else {
      pair_i[base + j] = static_cast<uint32_t>(v);
      pair_l[base + j] = static_cast<int32_t>(v >> 32);
    }
  }
}

__global__ __launch_bounds__(kLabelThreads) void k_label(
    Params* __restrict__ prm, const uint32_t* __restrict__ idx_a,
    const uint32_t* __restrict__ idx_b, const int32_t* __restrict__ point_cell,
    const uint8_t* __restrict__ is_core, const int32_t* __restrict__ cell_cluster,
    const uint32_t* __restrict__ spare_ia, const uint32_t* __restrict__ spare_ib,
    uint32_t* __restrict__ spare_xa, uint32_t* __restrict__ spare_xb,
    uint32_t* __restrict__ spare_ya, uint32_t* __restrict__ spare_yb,
    uint64_t* __restrict__ status, uint32_t* __restrict__ long_runs, int64_t n64, uint32_t B) {
  __shared__ uint32_t s_cnt[kMaxBuckets];  // counts, then the runs' local starts
  __shared__ uint64_t s_raw[kLabelTile];    // the tile's pairs, 64 KB in either form
  __shared__ uint32_t s_scan[kLabelThreads / 64];
  uint32_t* table = reinterpret_cast<uint32_t*>(status);
  if (prm->num_roots < kPackedRoots) {
    // the grid counts the two-word form's tiles; the packed form has half
    if (static_cast<int64_t>(blockIdx.x) * LabelForm<true>::kTile >= n64) return;
    label_tile<true>(prm, idx_a, idx_b, point_cell, is_core, cell_cluster, spare_ia, spare_ib,
                     spare_xa, spare_xb, spare_ya, spare_yb, table, long_runs, n64, B, s_cnt,
                     reinterpret_cast<uint32_t*>(s_raw), s_scan);
  } else {
    label_tile<false>(prm, idx_a, idx_b, point_cell, is_core, cell_cluster, spare_ia, spare_ib,
                      spare_xa, spare_xb, spare_ya, spare_yb, table, long_runs, n64, B, s_cnt,
                      s_raw, s_scan);
  }
}

// Inputs of more than kMaxBuckets * kUnpermMaxWidth points: labels written
// straight to their input positions (one 4-byte write per label).
__global__ __launch_bounds__(256) void k_label_direct(
    const Params* __restrict__ prm, const uint32_t* __restrict__ idx_a,
    const uint32_t* __restrict__ idx_b, const int32_t* __restrict__ point_cell,
    const uint8_t* __restrict__ is_core, const int32_t* __restrict__ cell_cluster,
    const uint32_t* __restrict__ spare_ia, const uint32_t* __restrict__ spare_ib,
    int32_t* __restrict__ labels, int64_t n, int32_t* __restrict__ error_out) {
  // the last kernel of a call: the error word is final (this kernel sets no bit)
  if (error_out && blockIdx.x == 0 && threadIdx.x == 0) error_out[0] = prm->error;
  // Workgroups go to the 8 XCDs in turn; each XCD takes a contiguous eighth of
  // the sorted points.  In a batch a cloud's points stay inside its own input
  // range, so one cloud's scattered label writes then gather in one XCD's L2
  // (whole lines leave it) instead of partial lines from all eight.
  const uint32_t g = gridDim.x, x = blockIdx.x & 7u, q8 = g >> 3, r8 = g & 7u;
  const uint32_t blk = x * q8 + min(x, r8) + (blockIdx.x >> 3);
  const int64_t i = static_cast<int64_t>(blk) * 256 + threadIdx.x;
  if (i >= n) return;
  const bool in_a = (prm->npasses & 1) != 0;
  const uint32_t dst = (in_a ? idx_a : idx_b)[i];
  const int32_t* __restrict__ lab_sorted =
      reinterpret_cast<const int32_t*>(in_a ? spare_ib : spare_ia);
  if (dst < static_cast<uint64_t>(n))
    labels[dst] = prm->error != 0 ? kPoison
                                  : (is_core[i] ? cell_cluster[point_cell[i]] : lab_sorted[i]);
}

// Pass B: one workgroup per bucket reads the bucket's run in every tile's
// slice (bounds from rows b and b+1 of the run table), places the labels in
// LDS by input position and writes them out in order (B <= kWidth).  Runs
// are mostly short (about B / tiles pairs), and then a group of 16 lanes
// reads kUnpermRuns runs at a time, their loads in flight together; the
// group's lane l holds, in registers, the bounds of the group's runs l,
// l + 16, ...  But the radix sort is stable, so a tile inside a dense cell
// holds its points in input order and gives a few buckets runs of hundreds
// of pairs, which would leave one group doing most of such a bucket: the
// label pass marks those buckets (a run of more than kLongRun pairs), and
// for them the work is split into chunks of 64 pairs, chunk k of run i going
// to wave (i + k) mod waves.  A wave then takes 64 runs at a time, one per
// lane, and walks its chunks of them in order: the owner of its j-th chunk
// is the last lane whose exclusive chunk count is <= j (a ballot), and
// kUnpermRuns chunks' loads are in flight together.
constexpr uint32_t kUnpermRuns = 8;
static_assert(16 % kUnpermRuns == 0, "a step's runs share a bound register");

template <bool kPacked>
__device__ __forceinline__ void place_label(int32_t* s_lab, uint32_t v, int32_t w, uint32_t b0,
                                            uint32_t b1) {
  if constexpr (kPacked) {
    const uint32_t d = b0 + (v & 0x7FFFu);
    if (d < b1) s_lab[d - b0] = static_cast<int32_t>(v >> 15) - 1;
  } else {
    if (v >= b0 && v < b1) s_lab[v - b0] = w;
  }
}

// kUnpermRuns runs of a 16-lane group: the group's runs r0 .. r0 + 7, tiles
// g + G r; lo / hi hold the bounds of run (r0 & ~15) + lane-in-group.
template <bool kPacked>
__device__ __forceinline__ void unpermute_runs(const uint32_t* __restrict__ pair_i,
                                               const int32_t* __restrict__ pair_l, int32_t* s_lab,
                                               uint32_t lo, uint32_t hi, uint32_t r0, uint32_t g,
                                               uint32_t gl, uint32_t G, uint32_t ntiles,
                                               uint32_t tile_sz, uint32_t n, uint32_t b0,
                                               uint32_t b1) {
  uint32_t s[kUnpermRuns], e[kUnpermRuns], off[kUnpermRuns];
#pragma unroll
  for (uint32_t u = 0; u < kUnpermRuns; ++u) {
    const uint32_t r = (r0 & 15u) + u;
    s[u] = __shfl(lo, static_cast<int>(r), 16);
    e[u] = __shfl(hi, static_cast<int>(r), 16);
    const uint32_t i = g + G * (r0 + u);
    off[u] = i * tile_sz;
    e[u] = i >= ntiles ? 0u : min(e[u], n - off[u]);
  }
  for (uint32_t c = gl;; c += 16u) {
    bool any = false;
    uint32_t v[kUnpermRuns];
    int32_t w[kUnpermRuns];
#pragma unroll
    for (uint32_t u = 0; u < kUnpermRuns; ++u) {
      if (s[u] + c < e[u]) {
        any = true;
        v[u] = pair_i[off[u] + s[u] + c];
        if constexpr (!kPacked) w[u] = pair_l[off[u] + s[u] + c];
      }
    }
    if (!any) break;
#pragma unroll
    for (uint32_t u = 0; u < kUnpermRuns; ++u)
      if (s[u] + c < e[u]) place_label<kPacked>(s_lab, v[u], w[u], b0, b1);
  }
}

// The chunks a wave takes of 64 runs (one per lane: its start, length, first
// chunk k0 for this wave and exclusive count of such chunks over the lanes).
template <bool kPacked>
__device__ __forceinline__ void unpermute_chunks(const uint32_t* __restrict__ pair_i,
                                                 const int32_t* __restrict__ pair_l,
                                                 int32_t* s_lab, uint32_t excl, uint32_t k0,
                                                 uint32_t start, uint32_t len, uint32_t total,
                                                 uint32_t lane, uint32_t W, uint32_t b0,
                                                 uint32_t b1) {
  for (uint32_t j = 0; j < total; j += kUnpermRuns) {
    uint32_t v[kUnpermRuns];
    int32_t w[kUnpermRuns];
    bool ok[kUnpermRuns];
#pragma unroll
    for (uint32_t u = 0; u < kUnpermRuns; ++u) {
      ok[u] = false;
      const uint32_t jj = j + u;
      if (jj >= total) continue;  // the same for the whole wave
      const uint64_t m = __ballot(excl <= jj);
      const int owner = 63 - __builtin_clzll(m);
      const uint32_t e_o = __builtin_amdgcn_readlane(excl, owner);
      const uint32_t k_o = __builtin_amdgcn_readlane(k0, owner);
      const uint32_t s_o = __builtin_amdgcn_readlane(start, owner);
      const uint32_t l_o = __builtin_amdgcn_readlane(len, owner);
      const uint32_t off = (k_o + (jj - e_o) * W) * 64u + lane;
      if (off < l_o) {
        ok[u] = true;
        v[u] = pair_i[s_o + off];
        if constexpr (!kPacked) w[u] = pair_l[s_o + off];
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < kUnpermRuns; ++u)
      if (ok[u]) place_label<kPacked>(s_lab, v[u], w[u], b0, b1);
  }
}

template <int kWidth, int kThreads>
__global__ __launch_bounds__(kThreads) void k_unpermute(
    const Params* __restrict__ prm, const uint32_t* __restrict__ spare_xa,
    const uint32_t* __restrict__ spare_xb, const uint32_t* __restrict__ spare_ya,
    const uint32_t* __restrict__ spare_yb, const uint32_t* __restrict__ table,
    const uint32_t* __restrict__ long_runs, int32_t* __restrict__ labels, int64_t n64,
    uint32_t B, int32_t* __restrict__ error_out) {
  // the last kernel of a call: the error word is final (this kernel sets no bit)
  if (error_out && blockIdx.x == 0 && threadIdx.x == 0) error_out[0] = prm->error;
  constexpr uint32_t W = kThreads / 64;  // waves
  constexpr uint32_t G = kThreads / 16;  // groups; group g reads the runs of tiles g, g + G, ...
  static_assert((W & (W - 1)) == 0, "waves per workgroup a power of two");
  // bound registers for the most tiles an input of this width has in the
  // two-word form (more tiles take the chunked path)
  constexpr uint32_t kMaxTiles =
      (static_cast<uint32_t>(kMaxBuckets) * kWidth + LabelForm<false>::kTile - 1) /
      LabelForm<false>::kTile;
  constexpr uint32_t kRL = (kMaxTiles + 16 * G - 1) / (16 * G);
  __shared__ int32_t s_lab[kWidth];
  const bool packed = prm->num_roots < kPackedRoots;
  const uint32_t n = static_cast<uint32_t>(n64);
  const uint32_t tile_sz = packed ? LabelForm<true>::kTile : LabelForm<false>::kTile;
  const uint32_t ntiles = (n + tile_sz - 1u) / tile_sz;
  const bool in_a = (prm->npasses & 1) != 0;
  const uint32_t* __restrict__ pair_i = in_a ? spare_xb : spare_xa;
  const int32_t* __restrict__ pair_l = reinterpret_cast<const int32_t*>(in_a ? spare_yb : spare_ya);
  const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
  // Workgroups go to the 8 XCDs in turn; each XCD takes a contiguous range of
  // buckets, so a cache line holding the ends of two neighbouring runs of a
  // tile is read into one XCD's L2, not two.
  const uint32_t nbk = gridDim.x, x = blockIdx.x & 7u, q8 = nbk >> 3, r8 = nbk & 7u;
  const uint32_t b = x * q8 + min(x, r8) + (blockIdx.x >> 3);
  const uint32_t b0 = b * B;
  const uint32_t b1 = min(n, b0 + min(B, static_cast<uint32_t>(kWidth)));
  const uint32_t* __restrict__ lo_row = table + static_cast<size_t>(b) * ntiles;
  const uint32_t* __restrict__ hi_row = lo_row + ntiles;
  if (long_runs[b] != 0u || ntiles > 16u * G * kRL) {
    for (uint32_t base = 0; base < ntiles; base += 64u) {
      const uint32_t i = base + lane;
      uint32_t start = 0u, len = 0u;
      if (i < ntiles) {
        const uint32_t lo = lo_row[i];
        const uint32_t hi = min(min(hi_row[i], tile_sz), n - i * tile_sz);
        start = i * tile_sz + lo;
        len = hi > lo ? hi - lo : 0u;
      }
      const uint32_t nch = (len + 63u) >> 6;
      const uint32_t k0 = (wv - i) & (W - 1u);  // the first chunk of run i this wave takes
      const uint32_t cnt = k0 < nch ? (nch - k0 + W - 1u) / W : 0u;
      uint32_t incl = cnt;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= static_cast<uint32_t>(d)) incl += o;
      }
      const uint32_t total = __shfl(incl, 63, 64);
      if (packed)
        unpermute_chunks<true>(pair_i, pair_l, s_lab, incl - cnt, k0, start, len, total, lane, W,
                               b0, b1);
      else
        unpermute_chunks<false>(pair_i, pair_l, s_lab, incl - cnt, k0, start, len, total, lane,
                                W, b0, b1);
    }
  } else {
    const uint32_t g = t >> 4, gl = t & 15u;
    const uint32_t g_wave = wv * 4u;  // the wave's first group
    uint32_t lo[kRL], hi[kRL];
#pragma unroll
    for (uint32_t q = 0; q < kRL; ++q) {
      const uint32_t i = g + G * (gl + 16u * q);
      lo[q] = hi[q] = 0u;
      if (i < ntiles) {
        lo[q] = lo_row[i];
        hi[q] = min(hi_row[i], tile_sz);
      }
    }
#pragma unroll
    for (uint32_t q = 0; q < kRL; ++q) {
      for (uint32_t r = 0; r < 16u; r += kUnpermRuns) {
        const uint32_t r0 = 16u * q + r;
        if (g_wave + G * r0 >= ntiles) break;  // the same for the whole wave
        if (packed)
          unpermute_runs<true>(pair_i, pair_l, s_lab, lo[q], hi[q], r0, g, gl, G, ntiles, tile_sz,
                               n, b0, b1);
        else
          unpermute_runs<false>(pair_i, pair_l, s_lab, lo[q], hi[q], r0, g, gl, G, ntiles,
                                tile_sz, n, b0, b1);
      }
    }
  }
  __syncthreads();
  // a call whose device-side checks failed leaves every label kPoison
  const bool failed = prm->error != 0;
  for (uint32_t j = b0 + t; j < b1; j += kThreads) labels[j] = failed ? kPoison : s_lab[j - b0];
}

// One window column [j0, j1) (column `col`, scan indices col * 5 + q) of a
// non-core point's border search by one lane.  The column's cells are
// visited by box distance, and a cell whose box is farther than the best d2
// so far cannot hold a better or tying core, so the rest of the column is
// skipped.  best / best_idx / lab keep the smallest (d2, scan index): the
// reference's first strict minimum in its scan order (dx-major, ky
// ascending), whatever the visiting order; all cores of a cell share its
// label.  Sets `deferred` when a big cell would need a scan.
__device__ __forceinline__ void lane_border_column(
    const float* __restrict__ xs, const float* __restrict__ ys,
    const uint8_t* __restrict__ is_core, float x, float y, int col, int32_t j0, int32_t j1,
    const int32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_flags,
    const uint4* __restrict__ cell_box, const int32_t* __restrict__ cell_cluster, float eps_sq,
    float& best, int32_t& best_idx, int32_t& lab, bool& deferred) {
  WinColumn w;
  load_column(cell_start, cell_flags, cell_box, j0, j1, w);
  float d[5];
  uint32_t todo = 0;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    d[q] = 0.f;
    if (q >= w.cnt || !(w.flags[q] & kCellHasCore)) continue;
    d[q] = point_box_d2(x, y, w.box[q]);
    if (d[q] > eps_sq) continue;
    todo |= 1u << q;
  }
#pragma unroll
  for (int it = 0; it < 5; ++it) {
    if (!todo) break;
    int qm = -1;
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (((todo >> q) & 1u) && (qm < 0 || d[q] < d[qm])) qm = q;
    todo &= ~(1u << qm);
    float dq = 0.f;
    int32_t cs = 0, ce = 0;
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (q == qm) {
        dq = d[q];
        cs = w.start[q];
        ce = w.start[q + 1];
      }
    if (dq > best) break;  // and every cell after it in this column
    if (ce - cs > kLaneScan) {
      deferred = true;
      continue;
    }
    const uint32_t cm = lane_min_core_d2(xs, ys, is_core, cs, ce, x, y, eps_sq);
    const int32_t idx = col * 5 + qm;
    if (cm != 0xFFFFFFFFu) {
      const float v = __uint_as_float(cm);
      if (v < best || (v == best && idx < best_idx)) {
        best = v;
        best_idx = idx;
        lab = cell_cluster[w.j0 + qm];
      }
    }
  }
}

// The same search by the whole wave for the point (xi, yi): each candidate
// cell is scanned 64 points per step; a step's smallest accepted d2 replaces
// the best when strictly smaller, so the first strict minimum in scan order
// wins as in the reference's sequential loop (cores of one cell share a
// label, so the order inside a cell does not matter).
__device__ int32_t wave_border(const float* __restrict__ xs, const float* __restrict__ ys,
                               const uint8_t* __restrict__ is_core, float xi, float yi,
                               const int32_t* __restrict__ row,
                               const int32_t* __restrict__ cell_start,
                               const uint32_t* __restrict__ cell_flags,
                               const uint4* __restrict__ cell_box,
                               const int32_t* __restrict__ cell_cluster, float eps_sq,
                               float best_init) {
  const int lane = threadIdx.x & 63;
  float best = best_init;
  int32_t lab = kNoise;
  for (int col = 0; col < 5; ++col) {
    const int32_t j0 = row[2 * col], j1 = row[2 * col + 1];
    for (int32_t nc = j0; nc < j1; ++nc) {
      if (!(cell_flags[nc] & kCellHasCore)) continue;
      if (point_box_d2(xi, yi, load_box(cell_box, nc)) > eps_sq) continue;
      const int32_t cs = cell_start[nc], ce = cell_start[nc + 1];
      for (int32_t base = cs; base < ce; base += 64) {
        const int32_t j = base + lane;
        uint32_t m = 0xFFFFFFFFu;
        if (j < ce && is_core[j]) {
          const float ex = xs[j] - xi, ey = ys[j] - yi;
          const float d2 = ex * ex + ey * ey;
          if (d2 <= eps_sq && d2 < best) m = __float_as_uint(d2);
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) m = min(m, static_cast<uint32_t>(__shfl_xor(m, d, 64)));
        if (m != 0xFFFFFFFFu) {
          best = __uint_as_float(m);
          lab = cell_cluster[nc];
        }
      }
    }
  }
  return lab;
}

// Border assignment (dbscan_grid.hpp:415-446) for the cells that hold a
// non-core point: a cell with no core cell in its window labels its points
// noise at once; the other cells' non-core points are searched 12 per step,
// five lanes per point (one window column each), and the label goes in
// sorted order into the index buffer the sort left spare (k_label moves it
// on).  Points whose search needs a big cell are searched by the whole wave.
template <class PP>
__global__ __launch_bounds__(256) void k_border(
    const Params* __restrict__ prm, const float* __restrict__ xs_a,
    const float* __restrict__ ys_a, uint32_t* __restrict__ spare_ia,
    const float* __restrict__ xs_b, const float* __restrict__ ys_b,
    uint32_t* __restrict__ spare_ib, const int32_t* __restrict__ cell_start,
    const int32_t* __restrict__ ngh, const uint32_t* __restrict__ cell_flags,
    const uint4* __restrict__ cell_box, const uint8_t* __restrict__ is_core,
    const int32_t* __restrict__ cell_cluster, PP pp) {
  const int64_t C = prm->num_cells;
  const bool in_a = (prm->npasses & 1) != 0;
  const float* __restrict__ xs = in_a ? xs_a : xs_b;
  const float* __restrict__ ys = in_a ? ys_a : ys_b;
  int32_t* __restrict__ lab_sorted = reinterpret_cast<int32_t*>(in_a ? spare_ib : spare_ia);
  const int lane = threadIdx.x & 63;
  const int64_t wave = (static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x) >> 6;
  const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * 256) >> 6;
  const int64_t K = cells_per_wave(C);
  for (int64_t base = wave * K; base < C; base += nwaves * K) {
    const int32_t c = static_cast<int32_t>(base + lane);
    const bool valid = lane < K && base + lane < C;
    int32_t s = 0, e = 0;
    bool has = false;
    CellParams cp{0.f, 0.f, 0};
    if (valid && (cell_flags[c] & kCellHasNonCore)) {
      has = true;
      s = cell_start[c];
      e = cell_start[c + 1];
      cp = pp.at(c);
      // no core cell in the 25-cell window: every point of the cell is noise
      // (two rounds of loads for the cell rather than a search per point)
      const int32_t* row = ngh + static_cast<int64_t>(c) * 10;
      int32_t j[10];
#pragma unroll
      for (int k = 0; k < 10; ++k) j[k] = row[k];
      uint32_t fl[25];
#pragma unroll
      for (int k = 0; k < 5; ++k)
#pragma unroll
        for (int q = 0; q < 5; ++q)
          fl[5 * k + q] = j[2 * k] + q < j[2 * k + 1] ? cell_flags[j[2 * k] + q] : 0u;
      uint32_t any = 0;
#pragma unroll
      for (int k = 0; k < 25; ++k) any |= fl[k];
      if (!(any & kCellHasCore)) {
        for (int32_t p = s; p < e; ++p) lab_sorted[p] = kNoise;
        has = false;
      }
    }
    const float eps_sq = cp.eps_sq, best_init = cp.best_init;
    // The points of these cells are searched 12 per step, each by 5 lanes,
    // one window column per lane (lanes 60-63 idle): point item q of the
    // wave is point q - excl of the cell of the lane whose [excl, incl)
    // holds q.  The five columns' loads go out together, and the lanes'
    // results are merged into the smallest (d2, scan index).  Core points
    // are skipped.
    const uint32_t cnt = has ? static_cast<uint32_t>(e - s) : 0u;
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
    const int grp = lane / 5, col = lane - 5 * grp;
    for (uint32_t r = 0; r < total; r += 12) {
      const uint32_t q = r + static_cast<uint32_t>(grp);
      int o = 0;  // owner lane: the number of lanes with incl <= q
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1) {
        const uint32_t v = static_cast<uint32_t>(__shfl(static_cast<int>(incl), o + step - 1, 64));
        if (v <= q) o += step;
      }
      o = min(o, 63);
      const int32_t oc = __shfl(c, o, 64), os = __shfl(s, o, 64);
      const uint32_t oex = static_cast<uint32_t>(__shfl(static_cast<int>(incl - cnt), o, 64));
      const float oeps = __shfl(eps_sq, o, 64), obest = __shfl(best_init, o, 64);
      bool deferred = false, work = false;
      int32_t p = 0, lab = kNoise, best_idx = 1 << 30;
      float best = obest;
      if (lane < 60 && q < total) {
        p = os + static_cast<int32_t>(q - oex);
        work = !is_core[p];
        if (work) {
          const int32_t* row = ngh + static_cast<int64_t>(oc) * 10;
          lane_border_column(xs, ys, is_core, xs[p], ys[p], col, row[2 * col], row[2 * col + 1],
                             cell_start, cell_flags, cell_box, cell_cluster, oeps, best,
                             best_idx, lab, deferred);
        }
      }
      // merge the point's five columns into its column-0 lane
#pragma unroll
      for (int dd = 1; dd <= 4; dd <<= 1) {
        const float bo = __shfl_down(best, dd, 64);
        const int32_t io = __shfl_down(best_idx, dd, 64), lo = __shfl_down(lab, dd, 64);
        const bool fo = __shfl_down(static_cast<int>(deferred), dd, 64) != 0;
        if (col + dd < 5) {
          if (bo < best || (bo == best && io < best_idx)) {
            best = bo;
            best_idx = io;
            lab = lo;
          }
          deferred = deferred || fo;
        }
      }
      const bool lead = work && col == 0;
      if (lead && !deferred) lab_sorted[p] = lab;
      // points whose search needs a big cell: the whole wave, one by one
      uint64_t dm = __ballot(lead && deferred);
      while (dm) {
        const int src = lowest_lane(dm);
        dm &= dm - 1ull;
        const int32_t pc = bcast(oc, src), pp_ = bcast(p, src);
        const int32_t* row = ngh + static_cast<int64_t>(pc) * 10;
        const int32_t wl = wave_border(xs, ys, is_core, xs[pp_], ys[pp_], row, cell_start,
                                       cell_flags, cell_box, cell_cluster, bcastf(oeps, src),
                                       bcastf(obest, src));
        if (lane == src) lab_sorted[pp_] = wl;
      }
    }
  }
}

// The error word of the call that last used the workspace, to dst (device
// memory or pinned host memory), in stream order after that call.
__global__ __launch_bounds__(64) void k_copy_error(const Params* __restrict__ prm,
                                                   int32_t* __restrict__ dst) {
  if (threadIdx.x == 0) dst[0] = prm->error;
}

// ---------------------------------------------------------------------------
// Small clouds: one workgroup clusters a whole cloud in LDS
// ---------------------------------------------------------------------------
//
// A cloud of at most kSmallMax points runs the reference's stages
// (dbscan_grid.hpp:159-454) inside one workgroup, its working set in LDS,
// with barriers between the stages instead of the pipeline's ~20 launches
// (which are most of a small call's time):
//   1. points into LDS, cell-coordinate bounds (block min/max);
//   2. stable LSD radix sort of 16-bit point indices by the compressed key
//      x' << by | y' (8-bit digits; wave64 ballot ranking into per-wave
//      counters, as in the pipeline's radix pass), keys recomputed from the
//      points each pass; the points are then gathered into sorted order;
//   3. cell runs by a block scan: point -> cell, cell starts and keys;
//   4. per cell: its 5x5 window as five (first cell, count) column ranges by
//      binary search over the cell keys, and its reference-order index (a
//      rotation inside its column when ky straddles 0, as in k_ngh);
//   5. core marking (dbscan_grid.hpp:267-307): a full cell wholesale, else the
//      own cell wholesale plus the distance test over the other 24 cells,
//      after the exact window-count prune;
//   6. union-find over reference indices in LDS (dbscan_grid.hpp:351-387): a
//      core point tests each later core window cell of another root for an
//      eps-close core point; CAS hooks the larger root onto the smaller;
//   7. cluster ids (dbscan_grid.hpp:392-398): roots ranked in reference order
//      by a block scan;
//   8. labels (dbscan_grid.hpp:403-454): a core point takes its cell's id, any
//      other the first strict minimum of (d2, window offset) over the cores
//      within eps; written to input order.
// Window columns are searched as k_ngh searches them: a neighbouring column
// outside the cloud's x' range holds no cell, and rows do not wrap (a cell
// coordinate is at most 2^31 - 128, so kx + 2 never wraps past INT32_MAX).
// Distances use the pipeline's test, fl(fl(ex*ex) + fl(ey*ey)) <= eps_sq.
constexpr int kSmallMax = DBSCAN_HIP_SMALL_MAX;
constexpr int kSmallThreads = 1024;
constexpr int kSmallWaves = kSmallThreads / 64;
constexpr int kSmallItems = kSmallMax / kSmallThreads;  // points per thread
constexpr int kSmallDigitBits = 8;
constexpr int kSmallDigits = 1 << kSmallDigitBits;
constexpr uint16_t kSmallNone = 0xFFFFu;  // cell without a cluster
// keys of at most this many bits are sorted by one counting pass
constexpr int kSmallCountBits = 14;
// clouds spanning fewer cell columns get a column table
constexpr int kSmallCols = 2048;
// a point whose window holds more points than this is done by a wave
constexpr int kSmallLaneScan = 64;
// lanes that test one cell pair together (union), one queued point (core)
constexpr int kSmallPairGroup = 8;
constexpr int kSmallCoreGroup = 16;
// clouds of at most this many cells get cell boxes for the union prunes
constexpr int kSmallBoxes = 1024;
// tests a lane makes on a cell pair before it queues the pair for a group
constexpr int kSmallLaneTests = 32;
static_assert(kSmallMax % kSmallThreads == 0, "whole items per thread");
static_assert(kSmallMax < 0xFFFF, "16-bit point and cell indices, 0xFFFF free");
static_assert(kSmallDigits <= kSmallThreads, "a thread per digit");

struct SmallLds {
  float x[kSmallMax], y[kSmallMax];  // input order, then sorted order
  uint16_t idx[2][kSmallMax];        // sort buffers; then sorted -> input, point -> cell
  uint8_t core[kSmallMax];           // by sorted point
  uint8_t has_core[kSmallMax];       // by cell
  uint16_t cstart[kSmallMax + 8];    // cell runs, C + 1 entries
  uint16_t cref[kSmallMax];          // reference-order index of each cell
  union {
    uint32_t wcnt[kSmallWaves][kSmallDigits];  // sort: per-wave digit counters
    uint64_t ckey[kSmallMax];                  // cells: compressed keys, ascending
    struct {
      uint32_t uf[kSmallMax];    // union-find parent, by reference index
      uint16_t rank[kSmallMax];  // root flags, then root ranks, by reference index
      uint16_t ccl[kSmallMax];   // cluster id by cell, or kSmallNone
    } u;
  } r;
  uint16_t wbase[kSmallMax][5];  // window column k (dx = k - 2): first cell
  uint16_t wlen[kSmallMax];      // its cell count (at most 5), 3 bits per column
  uint16_t col[kSmallCols + 8];  // first cell of column x' (narrow clouds)
  uint32_t dstart[kSmallDigits];
  uint32_t scan[kSmallWaves];
  int32_t red[kSmallWaves][4];
  int32_t err;
  int32_t ticket;
  int32_t qcore;   // points queued for a wave (core marking, then border)
};
static_assert(kSmallBoxes * sizeof(uint4) <= 2 * sizeof(uint16_t) * kSmallMax,
              "cell boxes in the rank and id space");
static_assert((1 << kSmallCountBits) / 2 * sizeof(uint32_t) <= sizeof(SmallLds::r),
              "16-bit counters, two to a word, in the cell-key space");
static_assert((1 << kSmallCountBits) / 2 % kSmallThreads == 0, "whole counter words per thread");
static_assert(kSmallMax < (1 << 16), "offsets and counts fit 16 bits");
static_assert(sizeof(SmallLds) <= 160 * 1024, "one workgroup's LDS on gfx950");

__device__ __forceinline__ uint32_t lds_ld(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_st(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// This lane's group of G lanes (G divides 64) in a wave ballot.
template <int G>
__device__ __forceinline__ uint64_t group_bits(uint64_t ballot) {
  if constexpr (G == 64) {
    return ballot;
  } else {
    const int base = static_cast<int>(threadIdx.x & 63) & ~(G - 1);
    return (ballot >> base) & ((uint64_t{1} << G) - 1u);
  }
}

// The o-th window column, nearest first (dx = 0, -1, 1, -2, 2).
__device__ __forceinline__ int near_column(int o) {
  return o == 0 ? 2 : ((o & 1) ? 2 - (o + 1) / 2 : 2 + o / 2);
}

// Root of x in the LDS forest, halving the path (parents only move to
// smaller indices, so every walk ends).  Bounded: a longer walk can only come
// from a corrupted forest (kErrUfLoop, stop there).
__device__ __forceinline__ uint32_t small_find(uint32_t* uf, uint32_t x, int32_t* err) {
  uint32_t p = lds_ld(uf + x);
  for (int it = 0; p != x; ++it) {
    if (it > kSmallMax) {
      atomicOr(err, static_cast<int32_t>(kErrUfLoop));
      break;
    }
    const uint32_t gp = lds_ld(uf + p);
    if (gp != p) lds_st(uf + x, gp);
    x = p;
    p = gp;
  }
  return x;
}

// A CAS fails only when another thread has just hooked that root, which
// happens at most C - 1 times, so 2 * kSmallMax attempts always suffice.
__device__ __forceinline__ void small_unite(uint32_t* uf, uint32_t a, uint32_t b, int32_t* err) {
  for (int it = 0; it < 2 * kSmallMax; ++it) {
    a = small_find(uf, a, err);
    b = small_find(uf, b, err);
    if (a == b) return;
    const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
    if (atomicCAS(uf + hi, hi, lo) == hi) return;
  }
  atomicOr(err, static_cast<int32_t>(kErrUfLoop));
}

// One cloud (ci: points [begin, end), parameters as a single call computes
// them), by the whole workgroup; returns after a closing barrier.
__device__ __forceinline__ void small_cloud(SmallLds& s, const CloudInfo& ci,
                                            const float2* __restrict__ X,
                                            int32_t* __restrict__ labels) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  // single-value stores are made by all of wave 0 (a wave-uniform branch)
  const bool wave0 = __builtin_amdgcn_readfirstlane(t) == 0;
  const int n = ci.end - ci.begin;
  if (n <= 0) return;
  if (n > kSmallMax) {  // the host routes only small clouds here
    if (wave0) atomicOr(&s.err, static_cast<int32_t>(kErrRange));
    return;
  }
  const float inv_r = ci.inv_r, eps_sq = ci.eps_sq;
  const int32_t ms = ci.ms;

  // 1. points and cell bounds
  int32_t b4[4] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  for (int i = t; i < n; i += kSmallThreads) {
    const float2 p = X[ci.begin + i];
    s.x[i] = p.x;
    s.y[i] = p.y;
    const int32_t kx = cell_coord(p.x * inv_r), ky = cell_coord(p.y * inv_r);
    b4[0] = min(b4[0], kx);
    b4[1] = max(b4[1], kx);
    b4[2] = min(b4[2], ky);
    b4[3] = max(b4[3], ky);
  }
  block_minmax4<kSmallWaves>(b4, s.red);  // ends with a barrier
  const uint32_t xmin = static_cast<uint32_t>(b4[0]), ymin = static_cast<uint32_t>(b4[2]);
  const uint32_t xspan = static_cast<uint32_t>(b4[1]) - xmin;  // x' <= xspan
  const uint32_t yspan = static_cast<uint32_t>(b4[3]) - ymin;  // y' <= yspan
  const uint32_t by = bit_width(yspan);
  const int nbits = static_cast<int>(bit_width(xspan) + by);
  const uint64_t ymask = (uint64_t{1} << by) - 1u;  // by <= 32
  auto key_of = [&](float x, float y) -> uint64_t {
    const uint32_t xr = static_cast<uint32_t>(cell_coord(x * inv_r)) - xmin;
    const uint32_t yr = static_cast<uint32_t>(cell_coord(y * inv_r)) - ymin;
    return (static_cast<uint64_t>(xr) << by) | yr;
  };

  // 2. point indices sorted by key.  No label depends on the order of a
  // cell's points, so a key of at most kSmallCountBits bits takes one
  // counting pass (LDS atomics); a wider one stable LSD passes of 8-bit digits.
  int src = 0;
  if (nbits == 0) {  // one cell
    for (int i = t; i < n; i += kSmallThreads) s.idx[0][i] = static_cast<uint16_t>(i);
  } else if (nbits <= kSmallCountBits) {
    // one 16-bit counter per key, two to a word (counts are at most 4,096)
    uint32_t* cnt = reinterpret_cast<uint32_t*>(&s.r);
    for (int e = t; e < (1 << kSmallCountBits) / 2; e += kSmallThreads) cnt[e] = 0u;
    __syncthreads();
    uint32_t key[kSmallItems], rk[kSmallItems];
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      const int i = t + k * kSmallThreads;
      key[k] = 0;
      rk[k] = 0;
      if (i < n) {
        key[k] = static_cast<uint32_t>(key_of(s.x[i], s.y[i]));
        const uint32_t sh = (key[k] & 1u) * 16u;
        rk[k] = (atomicAdd(cnt + (key[k] >> 1), 1u << sh) >> sh) & 0xFFFFu;
      }
    }
    __syncthreads();
    {
      constexpr int kWords = (1 << kSmallCountBits) / 2 / kSmallThreads;
      uint32_t w[kWords], sum = 0;
#pragma unroll
      for (int q = 0; q < kWords; ++q) {
        w[q] = cnt[t * kWords + q];
        sum += (w[q] & 0xFFFFu) + (w[q] >> 16);
      }
      uint32_t all;
      uint32_t off = block_excl_scan<kSmallThreads>(sum, all, s.scan);
#pragma unroll
      for (int q = 0; q < kWords; ++q) {
        const uint32_t lo = off, hi = off + (w[q] & 0xFFFFu);
        cnt[t * kWords + q] = lo | (hi << 16);
        off = hi + (w[q] >> 16);
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      const int i = t + k * kSmallThreads;
      if (i < n) {
        const uint32_t sh = (key[k] & 1u) * 16u;
        const uint32_t q = ((cnt[key[k] >> 1] >> sh) & 0xFFFFu) + rk[k];
        if (q < static_cast<uint32_t>(n)) s.idx[0][q] = static_cast<uint16_t>(i);
        else atomicOr(&s.err, static_cast<int32_t>(kErrRange));
      }
    }
  } else {
    // Item k of lane l in wave w sits at position w * 64 * kSmallItems +
    // k * 64 + l: ranking item by item keeps equal digits in position order.
    for (int i = t; i < n; i += kSmallThreads) s.idx[0][i] = static_cast<uint16_t>(i);
    const int passes = (nbits + kSmallDigitBits - 1) / kSmallDigitBits;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    for (int pass = 0; pass < passes; ++pass) {
      const uint32_t shift = static_cast<uint32_t>(pass * kSmallDigitBits);
      for (int e = t; e < kSmallWaves * kSmallDigits; e += kSmallThreads)
        (&s.r.wcnt[0][0])[e] = 0u;
      __syncthreads();
      uint32_t id[kSmallItems], dg[kSmallItems], rk[kSmallItems];
#pragma unroll
      for (int k = 0; k < kSmallItems; ++k) {
        const int p = wave * 64 * kSmallItems + k * 64 + lane;
        const bool valid = p < n;
        id[k] = valid ? s.idx[src][p] : 0u;
        const uint64_t key = valid ? key_of(s.x[id[k]], s.y[id[k]]) : 0ull;
        dg[k] = static_cast<uint32_t>(key >> shift) & (kSmallDigits - 1u);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kSmallDigitBits; ++b) {
          const bool bit = (dg[k] >> b) & 1u;
          const uint64_t bal = __ballot(bit);
          peers &= bit ? bal : ~bal;
        }
        const uint32_t below = static_cast<uint32_t>(__popcll(peers & lt_mask));
        uint32_t old = 0;
        if (valid) old = s.r.wcnt[wave][dg[k]];
        if (valid && below == 0)
          s.r.wcnt[wave][dg[k]] = old + static_cast<uint32_t>(__popcll(peers));
        rk[k] = old + below;
      }
      __syncthreads();
      // digit t: exclusive offsets of the waves, then of the digits
      uint32_t tot = 0;
      if (t < kSmallDigits) {
        for (int w = 0; w < kSmallWaves; ++w) {
          const uint32_t c = s.r.wcnt[w][t];
          s.r.wcnt[w][t] = tot;
          tot += c;
        }
      }
      uint32_t all;
      const uint32_t d0 = block_excl_scan<kSmallThreads>(tot, all, s.scan);
      if (t < kSmallDigits) s.dstart[t] = d0;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kSmallItems; ++k) {
        const int p = wave * 64 * kSmallItems + k * 64 + lane;
        if (p < n) {
          const uint32_t q = s.dstart[dg[k]] + s.r.wcnt[wave][dg[k]] + rk[k];
          if (q < static_cast<uint32_t>(n)) s.idx[src ^ 1][q] = static_cast<uint16_t>(id[k]);
          else atomicOr(&s.err, static_cast<int32_t>(kErrRange));
        }
      }
      src ^= 1;
      __syncthreads();
    }
  }
  __syncthreads();
  const uint16_t* __restrict__ perm = s.idx[src];  // sorted -> input
  uint16_t* __restrict__ pcell = s.idx[src ^ 1];   // sorted point -> cell
  {
    float gx[kSmallItems], gy[kSmallItems];
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      const int p = t + k * kSmallThreads;
      if (p < n) {
        const int i = perm[p];
        gx[k] = s.x[i];
        gy[k] = s.y[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      const int p = t + k * kSmallThreads;
      if (p < n) {
        s.x[p] = gx[k];
        s.y[p] = gy[k];
      }
    }
    __syncthreads();
  }

  // 3. cell runs (thread t: positions [t * kSmallItems, (t + 1) * kSmallItems))
  int C;
  {
    const int p0 = t * kSmallItems;
    uint64_t kk[kSmallItems];
    uint64_t prev = (p0 > 0 && p0 - 1 < n) ? key_of(s.x[p0 - 1], s.y[p0 - 1]) : 0ull;
    uint32_t flags = 0, cnt = 0;
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      const int p = p0 + k;
      kk[k] = 0;
      if (p < n) {
        kk[k] = key_of(s.x[p], s.y[p]);
        if (p == 0 || kk[k] != prev) {
          flags |= 1u << k;
          ++cnt;
        }
        prev = kk[k];
      }
    }
    uint32_t total;
    uint32_t c = block_excl_scan<kSmallThreads>(cnt, total, s.scan);
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      const int p = p0 + k;
      if (p < n) {
        if ((flags >> k) & 1u) {
          s.cstart[c] = static_cast<uint16_t>(p);
          s.r.ckey[c] = kk[k];
          ++c;
        }
        pcell[p] = static_cast<uint16_t>(c - 1);
      }
    }
    C = static_cast<int>(total);
    if (wave0) {
      s.cstart[C] = static_cast<uint16_t>(n);
      s.qcore = 0;
    }
  }
  __syncthreads();

  // 4. windows and reference indices.  Narrow clouds first get a column
  // table (col[x] = first cell of column x), so that every search below runs
  // inside one column; wider ones search all cells.
  const uint64_t* __restrict__ ckey = s.r.ckey;
  const bool coltab = xspan < static_cast<uint32_t>(kSmallCols);
  if (coltab) {
    for (int c = t; c < C; c += kSmallThreads) {
      const int xr = static_cast<int>(ckey[c] >> by);
      const int prev = c > 0 ? static_cast<int>(ckey[c - 1] >> by) : -1;
      for (int x = prev + 1; x <= xr; ++x) s.col[x] = static_cast<uint16_t>(c);
      if (c == C - 1)
        for (int x = xr + 1; x <= static_cast<int>(xspan) + 1; ++x)
          s.col[x] = static_cast<uint16_t>(C);
    }
    __syncthreads();
  }
  // first cell of [lo, lo + len) whose key is >= v
  auto lower_bound = [&](int lo, int len, uint64_t v) {
    while (len > 0) {
      const int half = len >> 1;
      if (ckey[lo + half] < v) {
        lo += half + 1;
        len -= half + 1;
      } else {
        len = half;
      }
    }
    return lo;
  };
  // cells of column x (x <= xspan): [lo, hi)
  auto column = [&](uint32_t x, int& lo, int& hi) {
    if (coltab) {
      lo = s.col[x];
      hi = s.col[x + 1];
    } else {
      lo = lower_bound(0, C, static_cast<uint64_t>(x) << by);
      hi = x < xspan ? lower_bound(lo, C - lo, static_cast<uint64_t>(x + 1u) << by) : C;
    }
  };
  const bool straddle = b4[2] < 0 && b4[3] >= 0;
  const uint32_t yr0 = 0u - ymin;  // y' of ky == 0 (when straddling)
  for (int c = t; c < C; c += kSmallThreads) {
    const uint64_t key = ckey[c];
    const uint32_t xr = static_cast<uint32_t>(key >> by);
    const uint32_t yr = static_cast<uint32_t>(key & ymask);
    const uint32_t ylo = yr >= 2u ? yr - 2u : 0u;
    const uint32_t yhi = static_cast<uint32_t>(
        min(static_cast<uint64_t>(yr) + 2u, static_cast<uint64_t>(yspan)));
    // the five lower bounds of (x' + dx, ylo), in lock-step, each inside its
    // column when there is a column table
    uint64_t lo_key[5], hi_key[5];
    int base[5], len[5];
    bool live[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const int64_t cx = static_cast<int64_t>(xr) + (k - 2);
      live[k] = cx >= 0 && cx <= static_cast<int64_t>(xspan);
      lo_key[k] = (static_cast<uint64_t>(cx) << by) | ylo;
      hi_key[k] = (static_cast<uint64_t>(cx) << by) | yhi;
      base[k] = 0;
      len[k] = 0;
      if (live[k]) {
        if (coltab) {
          base[k] = s.col[cx];
          len[k] = s.col[cx + 1] - base[k];
        } else {
          len[k] = C;
        }
      }
    }
    for (int it = 0; it < 16; ++it) {  // at most 13 halvings
      bool more = false;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        if (len[k] > 0) {
          const int half = len[k] >> 1;
          if (ckey[base[k] + half] < lo_key[k]) {
            base[k] += half + 1;
            len[k] -= half + 1;
          } else {
            len[k] = half;
          }
          more |= len[k] > 0;
        }
      }
      if (!more) break;
    }
    uint32_t packed = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      // the column's cells at y' in [ylo, yhi]: at most 5 (one per y'),
      // keys ascending, so a prefix from base (its length: the trailing
      // ones of the mask)
      uint32_t m = 0;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const int cq = min(base[k] + q, C - 1);
        m |= (live[k] && base[k] + q < C && ckey[cq] <= hi_key[k]) ? (1u << q) : 0u;
      }
      const int cnt = __builtin_ctz(~m);
      s.wbase[c][k] = static_cast<uint16_t>(live[k] ? base[k] : 0);
      packed |= static_cast<uint32_t>(cnt) << (3 * k);
    }
    s.wlen[c] = static_cast<uint16_t>(packed);
    int r = c;
    if (straddle) {
      int own_lo, own_hi;
      column(xr, own_lo, own_hi);
      const int split =
          lower_bound(own_lo, own_hi - own_lo, (static_cast<uint64_t>(xr) << by) | yr0);
      r = (yr >= yr0) ? own_lo + (c - split) : c + (own_hi - split);
    }
    if (r < 0 || r >= C) {
      atomicOr(&s.err, static_cast<int32_t>(kErrRange));
      r = c;
    }
    s.cref[c] = static_cast<uint16_t>(r);
    s.has_core[c] = 0;
  }
  __syncthreads();  // the cell keys are dead from here on (their space becomes uf)

  // The window of cell c as five ranges of sorted points, one per window
  // column (dx = k - 2): a column's window cells are consecutive, and so are
  // their points.  In scan order (dx, then dy, then the point) the ranges
  // and the points in them ascend.
  auto window = [&](int c, int (&lo)[5], int (&hi)[5]) {
    const uint32_t wl = s.wlen[c];
    int pop = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const int b = s.wbase[c][k];
      lo[k] = s.cstart[b];
      hi[k] = s.cstart[b + ((wl >> (3 * k)) & 7u)];
      pop += hi[k] - lo[k];
    }
    return pop;
  };
  auto dist_ok = [&](int j, float xi, float yi) {
    const float ex = s.x[j] - xi, ey = s.y[j] - yi;
    return ex * ex + ey * ey <= eps_sq;
  };

  // 5. core marking (dbscan_grid.hpp:267-307): a full cell wholesale; else
  // the own cell wholesale plus the eps-close points of the other 24 cells,
  // counted by the point's lane when its window is small and by a wave
  // (64 points a step) otherwise.
  uint32_t* __restrict__ uf = s.r.u.uf;
  uint16_t* __restrict__ pqueue = s.r.u.rank;  // free until stage 7
  for (int r = t; r < C; r += kSmallThreads) uf[r] = static_cast<uint32_t>(r);
  for (int p = t; p < n; p += kSmallThreads) {
    const int c = pcell[p];
    const int own_lo = s.cstart[c], own_hi = s.cstart[c + 1];
    const int size = own_hi - own_lo;
    bool core = size >= ms;  // dbscan_grid.hpp:271-279
    if (!core) {
      int lo[5], hi[5];
      const int pop = window(c, lo, hi);
      if (pop >= ms) {  // else fewer than ms points around: not core
        if (pop - size > kSmallLaneScan) {
          pqueue[atomicAdd(&s.qcore, 1)] = static_cast<uint16_t>(p);
          continue;
        }
        const float xi = s.x[p], yi = s.y[p];
        int count = size;
        for (int k = 0; k < 5; ++k)
          for (int j = lo[k]; j < hi[k]; ++j)
            count += (j < own_lo || j >= own_hi) && dist_ok(j, xi, yi) ? 1 : 0;
        core = count >= ms;
      }
    }
    s.core[p] = core ? 1 : 0;
    if (core) s.has_core[c] = 1;  // equal stores only
  }
  __syncthreads();
  {
    // groups of kSmallGroup lanes, a queued point each; the count is
    // order-free, so the nearest columns go first and it reaches ms sooner
    constexpr int G = kSmallCoreGroup;
    const int sub = lane & (G - 1), group = t / G;
    const int nq = __builtin_amdgcn_readfirstlane(s.qcore);
    for (int q = group; q < nq; q += kSmallThreads / G) {
      const int p = pqueue[q];
      const int c = pcell[p];
      const int own_lo = s.cstart[c], own_hi = s.cstart[c + 1];
      int lo[5], hi[5];
      window(c, lo, hi);
      const float xi = s.x[p], yi = s.y[p];
      int count = own_hi - own_lo;
      for (int o = 0; o < 5 && count < ms; ++o) {
        const int k = near_column(o);
        for (int j0 = lo[k]; j0 < hi[k] && count < ms; j0 += G) {
          const int j = j0 + sub;
          const bool hit = j < hi[k] && (j < own_lo || j >= own_hi) && dist_ok(j, xi, yi);
          count += static_cast<int>(__popcll(group_bits<G>(__ballot(hit))));
        }
      }
      const bool core = count >= ms;
      if (sub == 0) {
        s.core[p] = core ? 1 : 0;
        if (core) s.has_core[c] = 1;
      }
    }
  }
  __syncthreads();

  // 6. union-find over reference indices (dbscan_grid.hpp:351-387): each pair
  // of core cells in each other's window is tested once, from the lower
  // cell; CAS hooks the larger root onto the smaller.  A later window cell of
  // c is one of the next two of its own column or one of the (at most five)
  // of each column to its right.
  if (C > kSmallBoxes) {
    // many cells (a sparse cloud, few points per cell): a lane per cell
    for (int c = t; c < C; c += kSmallThreads) {
      if (!s.has_core[c]) continue;
      const uint32_t rc = s.cref[c];
      const int c_lo = s.cstart[c], c_hi = s.cstart[c + 1];
      const uint32_t wl = s.wlen[c];
      for (int k = 2; k < 5; ++k) {
        const int b = s.wbase[c][k], e = b + ((wl >> (3 * k)) & 7u);
        for (int nc = b > c ? b : c + 1; nc < e; ++nc) {
          if (!s.has_core[nc]) continue;
          const uint32_t rn = s.cref[nc];
          if (small_find(uf, rc, &s.err) == small_find(uf, rn, &s.err)) continue;
          const int n_lo = s.cstart[nc], n_hi = s.cstart[nc + 1];
          bool found = false;
          for (int i = c_lo; i < c_hi && !found; ++i) {
            if (!s.core[i]) continue;
            const float xi = s.x[i], yi = s.y[i];
            for (int j = n_lo; j < n_hi; ++j) {
              if (s.core[j] && dist_ok(j, xi, yi)) {
                found = true;
                break;
              }
            }
          }
          if (found) small_unite(uf, rc, rn, &s.err);
        }
      }
    }
    __syncthreads();
  } else {
    // Few cells, possibly large: the bounding box of each cell's finite core
    // points (ordered-integer bounds, in the rank and id space, free until
    // stage 7) prunes as the reference's boxes do (dbscan_grid.hpp:369-377):
    // a box pair farther apart than eps holds no eps-close pair, and a point
    // farther than eps from a box has no eps-close point in it (rounding is
    // monotone, so neither prune drops a pair the exact test would keep).
    uint4* __restrict__ box = reinterpret_cast<uint4*>(&s.r.u.rank[0]);
    for (int c = t; c < C; c += kSmallThreads)
      box[c] = make_uint4(0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u);
    __syncthreads();
    {
      // thread t: sorted points [t * kSmallItems, (t + 1) * kSmallItems),
      // one set of atomics per run of one cell
      const int p0 = t * kSmallItems;
      int cur = -1;
      uint4 acc = make_uint4(0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u);
      auto flush = [&]() {
        if (cur >= 0 && acc.x != 0xFFFFFFFFu) {
          atomicMin(&box[cur].x, acc.x);
          atomicMax(&box[cur].y, acc.y);
          atomicMin(&box[cur].z, acc.z);
          atomicMax(&box[cur].w, acc.w);
        }
      };
#pragma unroll
      for (int k = 0; k < kSmallItems; ++k) {
        const int p = p0 + k;
        if (p >= n) break;
        const int c = pcell[p];
        if (c != cur) {
          flush();
          cur = c;
          acc = make_uint4(0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u);
        }
        const float x = s.x[p], y = s.y[p];
        if (s.core[p] && isfinite(x) && isfinite(y)) {
          acc.x = min(acc.x, f2ord(x));
          acc.y = max(acc.y, f2ord(x));
          acc.z = min(acc.z, f2ord(y));
          acc.w = max(acc.w, f2ord(y));
        }
      }
      flush();
    }
    __syncthreads();
    auto box_of = [&](int c) {
      const uint4 u = box[c];
      return make_float4(ord2f(u.x), ord2f(u.y), ord2f(u.z), ord2f(u.w));
    };
    // pairs a lane leaves undecided go to groups of kSmallPairGroup lanes
    // (queued in the column table's space, dead now)
    uint32_t* __restrict__ pq = reinterpret_cast<uint32_t*>(&s.col[0]);
    constexpr int kPairQueue = (kSmallCols + 8) / 2;
    // Round 0 takes the pairs at most one column apart (nearly always joined
    // when both are core), round 1 the rest, most of which round 0 joined.
    for (int round = 0; round < 2; ++round) {
      if (wave0) s.qcore = 0;
      __syncthreads();
      const int slots = round == 0 ? 7 : 5;
      for (int item = t; item < slots * C; item += kSmallThreads) {
        const int c = item / slots, slot = item - slots * c;
        if (!s.has_core[c]) continue;
        int nc, e;
        if (round == 0 && slot < 2) {  // own column
          e = s.wbase[c][2] + ((s.wlen[c] >> 6) & 7u);
          nc = c + 1 + slot;
        } else {
          const int k = round == 0 ? 3 : 4;
          const int b = s.wbase[c][k];
          e = b + ((s.wlen[c] >> (3 * k)) & 7u);
          nc = b + (round == 0 ? slot - 2 : slot);
        }
        if (nc >= e || !s.has_core[nc]) continue;
        const float4 bn = box_of(nc);
        if (box_box_d2(box_of(c), bn) > eps_sq) continue;
        const uint32_t rc = s.cref[c], rn = s.cref[nc];
        if (small_find(uf, rc, &s.err) == small_find(uf, rn, &s.err)) continue;
        const int c_lo = s.cstart[c], c_hi = s.cstart[c + 1];
        const int n_lo = s.cstart[nc], n_hi = s.cstart[nc + 1];
        // The lane's share: at most kSmallLaneTests steps over (i, j), i
        // over c's points, j over nc's, branch-free (a point of c that is
        // not core or is farther than eps from nc's box skips its row); the
        // wave leaves the loop once none of its lanes has work left.
        bool found = false;
        int i = c_lo, j = n_lo;
        float xi = s.x[i], yi = s.y[i];
        bool row = s.core[i] && !(point_box_d2(xi, yi, bn) > eps_sq);
        for (int step = 0; step < kSmallLaneTests; step += 4) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const bool live = !found && i < c_hi;
            const int jc = j < n_hi ? j : n_hi - 1;
            const float ex = s.x[jc] - xi, ey = s.y[jc] - yi;
            found = found || (live && row && s.core[jc] && ex * ex + ey * ey <= eps_sq);
            const bool adv = live && !found && (!row || j + 1 >= n_hi);
            j = adv ? n_lo : (live && !found ? j + 1 : j);
            i = adv ? i + 1 : i;
            const int ic = i < c_hi ? i : c_hi - 1;
            const float nx = s.x[ic], ny = s.y[ic];
            const bool nrow = s.core[ic] && !(point_box_d2(nx, ny, bn) > eps_sq);
            xi = adv ? nx : xi;
            yi = adv ? ny : yi;
            row = adv ? nrow : row;
          }
          if (!__ballot(!found && i < c_hi)) break;
        }
        const bool done = found || i >= c_hi;
        if (found) {
          small_unite(uf, rc, rn, &s.err);
        } else if (!done) {
          const int q = atomicAdd(&s.qcore, 1);
          if (q < kPairQueue) {
            pq[q] = (static_cast<uint32_t>(c) << 16) | static_cast<uint32_t>(nc);
          } else {  // queue full: the lane finishes the pair
            for (int i = c_lo; i < c_hi && !found; ++i) {
              if (!s.core[i]) continue;
              const float xi = s.x[i], yi = s.y[i];
              if (point_box_d2(xi, yi, bn) > eps_sq) continue;
              for (int j = n_lo; j < n_hi; ++j) {
                if (s.core[j] && dist_ok(j, xi, yi)) {
                  found = true;
                  break;
                }
              }
            }
            if (found) small_unite(uf, rc, rn, &s.err);
          }
        }
      }
      __syncthreads();
      constexpr int G = kSmallPairGroup;
      const int sub = lane & (G - 1), group = t / G;
      const int nq = min(__builtin_amdgcn_readfirstlane(s.qcore), kPairQueue);
      for (int q = group; q < nq; q += kSmallThreads / G) {
        const uint32_t v = pq[q];
        const int c = static_cast<int>(v >> 16), nc = static_cast<int>(v & 0xFFFFu);
        const uint32_t rc = s.cref[c], rn = s.cref[nc];
        if (small_find(uf, rc, &s.err) == small_find(uf, rn, &s.err)) continue;
        const float4 bn = box_of(nc);
        const int c_lo = s.cstart[c], c_hi = s.cstart[c + 1];
        const int n_lo = s.cstart[nc], n_hi = s.cstart[nc + 1];
        bool found = false;
        for (int i0 = c_lo; i0 < c_hi && !found; i0 += G) {
          const int i = i0 + sub;
          bool act = i < c_hi && s.core[i];
          const float xi = act ? s.x[i] : 0.f, yi = act ? s.y[i] : 0.f;
          act = act && !(point_box_d2(xi, yi, bn) > eps_sq);
          if (!group_bits<G>(__ballot(act))) continue;
          for (int j = n_lo; j < n_hi; ++j) {
            if (!s.core[j]) continue;  // the same j in every lane of the group
            if (group_bits<G>(__ballot(act && dist_ok(j, xi, yi)))) {
              found = true;
              break;
            }
          }
        }
        if (found && sub == 0) small_unite(uf, rc, rn, &s.err);
      }
      __syncthreads();
    }
  }

  // 7. cluster ids: every reference index to its root, roots ranked in
  // reference order
  {
    uint32_t root[kSmallItems];
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      const int r = t + k * kSmallThreads;
      root[k] = r < C ? small_find(uf, static_cast<uint32_t>(r), &s.err) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      const int r = t + k * kSmallThreads;
      if (r < C) uf[r] = root[k];
    }
    __syncthreads();
  }
  for (int c = t; c < C; c += kSmallThreads) {
    const uint32_t r = s.cref[c];
    s.r.u.rank[r] = (s.has_core[c] && uf[r] == r) ? 1 : 0;
  }
  __syncthreads();
  {
    const int r0 = t * kSmallItems;
    uint32_t f[kSmallItems], sum = 0;
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      f[k] = r0 + k < C ? s.r.u.rank[r0 + k] : 0u;
      sum += f[k];
    }
    uint32_t roots;
    uint32_t off = block_excl_scan<kSmallThreads>(sum, roots, s.scan);
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      if (r0 + k < C) s.r.u.rank[r0 + k] = static_cast<uint16_t>(off);
      off += f[k];
    }
  }
  __syncthreads();
  for (int c = t; c < C; c += kSmallThreads)
    s.r.u.ccl[c] = s.has_core[c] ? s.r.u.rank[uf[s.cref[c]]] : kSmallNone;
  if (wave0) s.qcore = 0;
  __syncthreads();

  // 8. labels, in input order (dbscan_grid.hpp:403-454).  A core point takes
  // its cell's id; another the first strict minimum of d2 over the cores
  // within eps in scan order, i.e. the least (d2, sorted index) (both ascend
  // with the scan), by its lane for a small window, else by a wave.
  uint16_t* __restrict__ bqueue = s.r.u.rank;  // ranks are dead now
  const float best_init = ci.best_init;
  auto store = [&](int p, int32_t lab) {
    const int i = perm[p];
    if (i < n) labels[ci.begin + i] = lab;
    else atomicOr(&s.err, static_cast<int32_t>(kErrRange));
  };
  for (int p = t; p < n; p += kSmallThreads) {
    const int c = pcell[p];
    if (s.core[p]) {
      store(p, s.r.u.ccl[c]);
      continue;
    }
    int lo[5], hi[5];
    if (window(c, lo, hi) > kSmallLaneScan) {
      bqueue[atomicAdd(&s.qcore, 1)] = static_cast<uint16_t>(p);
      continue;
    }
    const float xi = s.x[p], yi = s.y[p];
    float best = best_init;
    int32_t lab = kNoise;
    for (int k = 0; k < 5; ++k) {
      for (int j = lo[k]; j < hi[k]; ++j) {
        if (!s.core[j]) continue;
        const float ex = s.x[j] - xi, ey = s.y[j] - yi;
        const float d2 = ex * ex + ey * ey;
        if (d2 <= eps_sq && d2 < best) {
          best = d2;
          lab = s.r.u.ccl[pcell[j]];
        }
      }
    }
    store(p, lab);
  }
  __syncthreads();
  {
    const int nq = __builtin_amdgcn_readfirstlane(s.qcore);
    for (int q = wave; q < nq; q += kSmallWaves) {
      const int p = __builtin_amdgcn_readfirstlane(bqueue[q]);
      int lo[5], hi[5];
      window(pcell[p], lo, hi);
      const float xi = s.x[p], yi = s.y[p];
      float best = best_init;
      int jbest = INT_MAX;
      for (int k = 0; k < 5; ++k) {
        for (int j = lo[k] + lane; j < hi[k]; j += 64) {
          if (!s.core[j]) continue;
          const float ex = s.x[j] - xi, ey = s.y[j] - yi;
          const float d2 = ex * ex + ey * ey;
          if (d2 <= eps_sq && d2 < best) {  // a lane's j ascend: the first wins ties
            best = d2;
            jbest = j;
          }
        }
      }
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) {
        const float ob = __shfl_xor(best, d, 64);
        const int oj = __shfl_xor(jbest, d, 64);
        if (ob < best || (ob == best && oj < jbest)) {
          best = ob;
          jbest = oj;
        }
      }
      jbest = __builtin_amdgcn_readfirstlane(jbest);
      if (lane == 0) store(p, jbest == INT_MAX ? kNoise : static_cast<int32_t>(s.r.u.ccl[pcell[jbest]]));
    }
  }
  __syncthreads();
}

// Clears the batch-mode words of k_small (ticket, finished count, error word).
__global__ __launch_bounds__(64) void k_small_init(Params* __restrict__ prm, int32_t error0) {
  if (threadIdx.x == 0) {
    prm->error = error0;
    prm->tile_counter[0] = 0u;
    prm->tile_counter[1] = 0u;
  }
}

// Clusters num_clouds small clouds, one workgroup at a time per cloud
// (clouds == nullptr: the single cloud `one`).  One workgroup walks the
// clouds in order and owns the error word (error0 is where it starts).  A
// larger grid is persistent: its workgroups take clouds from an atomic
// ticket (cleared by k_small_init, which also set the error word), OR their
// errors into the word, and the last one to finish reads it.  Either way a
// call whose word ends non-zero has every one of its `total` labels
// overwritten with kPoison, and the word goes to error_out (if not null).
__global__ __launch_bounds__(kSmallThreads) void k_small(
    const float2* __restrict__ X, int32_t* __restrict__ labels,
    const CloudInfo* __restrict__ clouds, int32_t num_clouds, CloudInfo one,
    Params* __restrict__ prm, int32_t error0, int32_t* __restrict__ error_out, int64_t total) {
  __shared__ SmallLds s;
  // branches on these are uniform across a wave (see k_sort_tail)
  const bool wave0 = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x)) == 0;
  const uint32_t one_inc = threadIdx.x == 0 ? 1u : 0u;
  if (wave0) s.err = 0;
  __syncthreads();
  for (int32_t k = -1;;) {
    if (gridDim.x == 1) {
      ++k;
    } else {
      if (wave0) {
        const uint32_t old = atomicAdd(&prm->tile_counter[0], one_inc);
        s.ticket = __builtin_amdgcn_readlane(static_cast<int>(old), 0);
      }
      __syncthreads();
      k = __builtin_amdgcn_readfirstlane(s.ticket);
      __syncthreads();
    }
    if (k >= num_clouds) break;
    small_cloud(s, clouds ? clouds[k] : one, X, labels);
  }
  __syncthreads();
  const int32_t local = __builtin_amdgcn_readfirstlane(s.err);
  int32_t bits;
  if (gridDim.x == 1) {
    bits = local | error0;
    if (wave0) prm->error = bits;
  } else {
    if (local && wave0) atomicOr(&prm->error, local);
    __threadfence();
    __syncthreads();
    if (wave0) {
      const uint32_t old = atomicAdd(&prm->tile_counter[1], one_inc);
      s.ticket = __builtin_amdgcn_readlane(static_cast<int>(old), 0);
    }
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(s.ticket) != static_cast<int32_t>(gridDim.x) - 1) return;
    __threadfence();
    bits = __builtin_amdgcn_readfirstlane(ld_relaxed(&prm->error));
  }
  if (bits) {
    for (int64_t i = threadIdx.x; i < total; i += kSmallThreads) labels[i] = kPoison;
  }
  if (error_out && wave0) error_out[0] = bits;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Workspace layout for n points.  clouds > 0: a batch of that many clouds
// holding n points in all (tile counts are bounded by n / tile + clouds).
struct Layout {
  size_t params, sort_status, cell_status, rank_status, cloud_hist, long_runs, zero_end;
  size_t partials, clouds, seg_sort, seg_scan, seg_end;
  size_t xs_a, ys_a, idx_a, xs_b, ys_b, idx_b;
  size_t point_cell, is_core, cell_start, cell_key, cell_flags, cell_box, ngh, cell_ref, uf, rank,
      cell_cluster, col_range, cell_cloud;
  int64_t max_cols;
  size_t total;
};

constexpr size_t align256(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }

Layout make_layout(int64_t n, int64_t clouds = 0) {
  const size_t N = static_cast<size_t>(n > 0 ? n : 1);
  const size_t B = static_cast<size_t>(clouds > 0 ? clouds : 0);
  const size_t sort_tiles = (N + kSortTileMin - 1) / kSortTileMin + B;
  const size_t scan_tiles = (N + kScanTile - 1) / kScanTile + B;
  Layout L{};
  size_t off = 0;
  auto take = [&off](size_t bytes) {
    const size_t at = off;
    off = align256(off + bytes);
    return at;
  };
  L.params = take(sizeof(Params));
  // the radix passes' look-back words, reused for the label pass's run table
  // ((kMaxBuckets + 1) x tiles words)
  const size_t label_tiles = (N + kLabelTile - 1) / kLabelTile;
  L.sort_status = take(std::max(sort_tiles * kRadixMax, label_tiles * kMaxBuckets) *
                       sizeof(uint64_t));
  L.cell_status = take(scan_tiles * sizeof(uint64_t));
  L.rank_status = take(scan_tiles * sizeof(uint64_t));
  L.cloud_hist = take(B * kMaxPasses * kRadixMax * sizeof(uint32_t));
  L.long_runs = take(kMaxBuckets * sizeof(uint32_t));
  L.zero_end = off;
  L.partials = take(kMaxReduceBlocks * sizeof(int4));
  L.clouds = take(B * sizeof(CloudInfo));
  L.seg_sort = take(B ? sort_tiles * sizeof(SegTile) : 0);
  L.seg_scan = take(B ? scan_tiles * sizeof(SegTile) : 0);
  L.seg_end = off;
  L.xs_a = take(N * 4);
  L.ys_a = take(N * 4);
  L.idx_a = take(N * 4);
  L.xs_b = take(N * 4);
  L.ys_b = take(N * 4);
  L.idx_b = take(N * 4);
  L.point_cell = take(N * 4);
  L.is_core = take(N);
  L.cell_start = take((N + 1) * 4);
  L.cell_key = take(N * 8);
  L.cell_flags = take(N * 4);
  L.cell_box = take(N * 16);
  L.ngh = take(N * 10 * 4);
  L.cell_ref = take(N * 4);
  L.uf = take(N * 4);
  L.rank = take(N * 4);
  L.cell_cluster = take(N * 4);
  L.max_cols = static_cast<int64_t>(std::min<size_t>(size_t(1) << 22, 4 * N + 64));
  L.col_range = take(static_cast<size_t>(L.max_cols) * sizeof(int2));
  L.cell_cloud = take(B ? N * 4 : 0);
  L.total = off;
  return L;
}

thread_local int g_last_hip_error = 0;
// Error bits the next call on this host thread starts with (test hook:
// dbscan2d_hip_inject_error_for_testing); consumed by that call.
thread_local int32_t g_injected_error = 0;

int32_t take_injected_error() {
  const int32_t v = g_injected_error;
  g_injected_error = 0;
  return v;
}

const char* const kStageNames[DBSCAN_HIP_NUM_STAGES] = {
    "cell_keys", "sort_gather", "boundary_scan", "build_ngh", "core_mark",
    "union_find", "cluster_id", "core_labels", "border_assign", "unpermute"};

// Bucket width B of the label unpermute (input positions per bucket), or 0
// when more than kMaxBuckets buckets of kUnpermMaxWidth would be needed
// (k_label_direct then writes the labels).
uint32_t bucket_width(int64_t n) {
  int64_t B = (n + kMaxBuckets - 1) / kMaxBuckets;
  B = std::max<int64_t>(4096, (B + 255) / 256 * 256);
  return B <= kUnpermMaxWidth ? static_cast<uint32_t>(B) : 0u;
}

void launch_unpermute(uint32_t B, int64_t n, hipStream_t stream, const Params* prm,
                      float* xs_a, float* xs_b, float* ys_a, float* ys_b, const uint64_t* status,
                      const uint32_t* long_runs, int32_t* labels, int32_t* error_out) {
  const int nb = static_cast<int>((n + B - 1) / B);
  auto u = [](float* p) { return reinterpret_cast<uint32_t*>(p); };
  const uint32_t* table = reinterpret_cast<const uint32_t*>(status);
  if (B <= static_cast<uint32_t>(kUnpermSmallWidth))
    k_unpermute<kUnpermSmallWidth, 512><<<nb, 512, 0, stream>>>(
        prm, u(xs_a), u(xs_b), u(ys_a), u(ys_b), table, long_runs, labels, n, B, error_out);
  else
    k_unpermute<kUnpermMaxWidth, 1024><<<nb, 1024, 0, stream>>>(
        prm, u(xs_a), u(xs_b), u(ys_a), u(ys_b), table, long_runs, labels, n, B, error_out);
}

int grid_for(int64_t items, int per_block, int cap);

// The label pass (its form chosen on the device).
void launch_label(uint32_t B, int64_t n, hipStream_t stream, Params* prm, const uint32_t* idx_a,
                  const uint32_t* idx_b, const int32_t* point_cell, const uint8_t* is_core,
                  const int32_t* cell_cluster, float* xs_a, float* xs_b, float* ys_a,
                  float* ys_b, uint64_t* status, uint32_t* long_runs) {
  auto u = [](float* p) { return reinterpret_cast<uint32_t*>(p); };
  k_label<<<grid_for(n, kLabelTile, INT_MAX), kLabelThreads, 0, stream>>>(
      prm, idx_a, idx_b, point_cell, is_core, cell_cluster, idx_a, idx_b, u(xs_a), u(xs_b),
      u(ys_a), u(ys_b), status, long_runs, n, B);
}

int grid_for(int64_t items, int per_block, int cap) {
  int64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<int>(g);
}

template <typename T>
T* at(void* ws, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(ws) + off);
}

// Launch order of the kernels; dbscan2d_hip_timed records an event before
// each of them and one after the last.
enum KernelSlot : int {
  kKBounds = 0,
  kKHist,
  kKSort0,
  kKCells = kKSort0 + kMaxPasses,
  kKBbox,
  kKNgh,
  kKCore,
  kKUnion,
  kKUnionLocal,
  kKFlatten0,
  kKUnion1,
  kKRank,
  kKCluster,
  kKBorder,
  kKLabel,
  kKUnpermute,
  kNumKernels
};
static_assert(kNumKernels == DBSCAN_HIP_NUM_KERNELS, "kernel count in the header");

const char* const kKernelNames[kNumKernels] = {
    "k_bounds", "k_hist", "k_sort_pass[0]", "k_sort_pass[1]", "k_sort_pass[2]",
    "k_sort_pass[3]", "k_sort_pass[4]", "k_sort_pass[5]", "k_sort_pass[6]", "k_sort_pass[7]",
    "k_cells", "k_bbox", "k_ngh", "k_core_sparse", "k_union[3x3]", "k_union_local", "k_flatten", "k_union[rest]", "k_rank", "k_cluster", "k_border", "k_label", "k_unpermute"};

// Kernel slots [begin, end) of each reference stage.  Border assignment runs
// before the core labels here: it writes sorted-order labels that the label
// pass (k_label: pairs into buckets) then moves with the core ones.
constexpr int kStageSlots[DBSCAN_HIP_NUM_STAGES][2] = {
    {kKBounds, kKSort0}, {kKSort0, kKCells},   {kKCells, kKNgh},        {kKNgh, kKCore},
    {kKCore, kKUnion},   {kKUnion, kKRank},    {kKRank, kKBorder},      {kKLabel, kKUnpermute},
    {kKBorder, kKLabel}, {kKUnpermute, kNumKernels}};

// Built-in launch configuration by input size: the autotuner's table for
// MI355X (dbscan_torch/tuning_gfx950.json); a caller can pass its own.
// (Widening radix_bits changes nothing unless it saves a pass: the digit
// width is the key width over the pass count.  At 10M points, 9 bits sort
// 25-27-bit keys in 3 passes of about 94 us instead of 4 of about 74 us.)
// An all-small batch runs as one persistent grid while it has at most this
// many clouds per CU.
constexpr int kSmallGridClouds = 6;

// Clouds of at most kSmallDefault points take the one-workgroup path: on one
// MI355X it beats the pipeline at every size it takes (single calls of
// constant-density clouds: 45 against 113 us at 100 points, 80 against 136
// us at 4,096; batches of 16 to 1,024 such clouds 1.4-2x).
constexpr int kSmallDefault = kSmallMax;

dbscan_hip_config default_config(int64_t n) {
  if (n <= 31622) return dbscan_hip_config{8, 8, 2048, 16384, kSmallDefault};
  if (n <= 316227) return dbscan_hip_config{8, 9, 16384, 4096, kSmallDefault};
  if (n <= 2000000) return dbscan_hip_config{8, 9, 8192, 8192, kSmallDefault};
  if (n <= 3162277) return dbscan_hip_config{12, 9, 8192, 8192, kSmallDefault};
  if (n <= (int64_t{1} << 24)) return dbscan_hip_config{12, 9, 4096, 8192, kSmallDefault};
  // mostly batches of many clouds: 32,768 workgroups for the per-cell and
  // pair kernels (the 135M-point 2,048-cloud batch 11.56 -> 11.30 ms against
  // 8,192; a single 20M-point call unchanged)
  return dbscan_hip_config{12, 9, 32768, 32768, kSmallDefault};
}

bool config_ok(const dbscan_hip_config& c) {
  return (c.sort_items == 8 || c.sort_items == 12 || c.sort_items == 16) && c.radix_bits >= 4 &&
         c.radix_bits <= kRadixBitsMax && c.cell_blocks >= 64 && c.cell_blocks <= 65536 &&
         c.pair_blocks >= 64 && c.pair_blocks <= 65536 && c.small_max >= 0 &&
         c.small_max <= kSmallMax;
}

// Passes from kTailPass on are needed only for keys wider than 4 digits of
// the configuration's radix_bits (36 bits at 9); they run in one launch
// (k_sort_tail) whose workgroups claim tiles by ticket, so that for the usual
// narrow key they cost a single launch that returns at once.  Its grid is two
// workgroups per CU (any size is correct: no workgroup waits for one that has
// not claimed a tile).  The CU count is read once per device.
constexpr int kMaxDevices = 64;
std::atomic<int> g_cu_count[kMaxDevices];

int cu_count() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 1;
  if (dev < kMaxDevices) {
    const int c = g_cu_count[dev].load(std::memory_order_relaxed);
    if (c > 0) return c;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    return 1;
  if (dev < kMaxDevices) g_cu_count[dev].store(cus, std::memory_order_relaxed);
  return cus;
}

int tail_grid(int tiles) { return std::max(1, std::min(tiles, 2 * cu_count())); }

// Radix pass `pass` with `items` points per thread over `tiles` tiles.
template <bool kFirst, bool kSeg, class... A>
void launch_sort(int items, int tiles, hipStream_t st, A... a) {
  const uint32_t nt = static_cast<uint32_t>(tiles);
  if (items == 8)
    k_sort_pass<kFirst, kSeg, 8><<<tiles, kSortThreads, 0, st>>>(a..., nt);
  else if (items == 16)
    k_sort_pass<kFirst, kSeg, 16><<<tiles, kSortThreads, 0, st>>>(a..., nt);
  else
    k_sort_pass<kFirst, kSeg, 12><<<tiles, kSortThreads, 0, st>>>(a..., nt);
}

// Passes first_pass.. in one launch (k_sort_tail).
template <bool kSeg, class... A>
void launch_sort_tail(int items, int tiles, int first_pass, hipStream_t st, A... a) {
  const uint32_t nt = static_cast<uint32_t>(tiles);
  const int g = tail_grid(tiles);
  if (items == 8)
    k_sort_tail<kSeg, 8><<<g, kSortThreads, 0, st>>>(a..., first_pass, nt);
  else if (items == 16)
    k_sort_tail<kSeg, 16><<<g, kSortThreads, 0, st>>>(a..., first_pass, nt);
  else
    k_sort_tail<kSeg, 12><<<g, kSortThreads, 0, st>>>(a..., first_pass, nt);
}

// Inputs of at most this many radix tiles (160k points at 4,096-point
// tiles) run pass 1 as a plain launch and passes 2.. in the tail launch (a
// kernel boundary costs more than handing a pass over inside the launch
// when the passes are this short: 100k points, 25 tiles, 4.5% faster per
// call than with a limit of 16 tiles; at 61 tiles the hand-off was slower).  Pass 1 stays a plain launch because a
// small cloud's key usually needs only two passes, and a plain pass is
// cheaper than the tail's ticketed one while a tail with no pass to run
// costs nothing measurable (10k-60k points 1-5% faster per call on one
// MI355X than the tail from pass 1).  Inputs of at most kDirectLabelMax
// points write their labels directly (k_label_direct) instead of through the
// two bucketed passes: one launch instead of two, and up to about 1.5M points
// the scattered 4-byte writes cost less than the second pass (300k points 9%
// faster per call, 1.07M points 3%; at 10M points the direct write is 21%
// slower).
constexpr int kTailFromOneTiles = 40;
constexpr int64_t kDirectLabelMax = int64_t{3} << 19;  // 1,572,864
int tail_first_pass(int tiles) { return tiles <= kTailFromOneTiles ? 2 : kTailPass; }
uint32_t label_buckets(int64_t n) { return n <= kDirectLabelMax ? 0u : bucket_width(n); }

int run(const float* X, int64_t n, double eps, int64_t min_samples64, int32_t* labels,
        void* workspace, size_t workspace_bytes, hipStream_t stream, float* stage_ms,
        hipEvent_t* kev, const dbscan_hip_config* cfgp, int32_t* error_out = nullptr) {
  if (n < 0 || n > INT32_MAX) return DBSCAN_HIP_ERR_N;
  if (!(eps > 0.0)) return DBSCAN_HIP_ERR_EPS;  // csrc/extension.cpp:29
  if (min_samples64 < 1) return DBSCAN_HIP_ERR_MIN_SAMPLES;  // csrc/extension.cpp:30
  if (stage_ms)
    for (int s = 0; s < DBSCAN_HIP_NUM_STAGES; ++s) stage_ms[s] = 0.f;
  if (n == 0) return DBSCAN_HIP_OK;
  if (!X || !labels || !workspace) return DBSCAN_HIP_ERR_NULL;
  if ((reinterpret_cast<uintptr_t>(X) & 7u) || (reinterpret_cast<uintptr_t>(labels) & 3u) ||
      (reinterpret_cast<uintptr_t>(workspace) & 255u))
    return DBSCAN_HIP_ERR_ALIGN;
  const Layout L = make_layout(n);
  if (workspace_bytes < L.total) return DBSCAN_HIP_ERR_WORKSPACE;
  const dbscan_hip_config cfg = cfgp ? *cfgp : default_config(n);
  if (!config_ok(cfg)) return DBSCAN_HIP_ERR_CONFIG;

  // Scalar parameters exactly as the reference computes them
  // (dbscan_cpu.cpp:52-53, dbscan_grid.hpp:153-154,425).
  const float eps_f = static_cast<float>(eps);
  const int32_t ms = static_cast<int32_t>(static_cast<uint32_t>(static_cast<uint64_t>(min_samples64)));
  const float inv_r = static_cast<float>(std::sqrt(2.0)) / eps_f;
  const float eps_sq = eps_f * eps_f;
  const float best_init = eps_sq + 1.0f;
  const OneCloud pp{eps_sq, best_init, ms};

  hipEvent_t ev[kNumKernels + 1] = {};  // profile: one before every kernel slot
  auto release = [&]() {
    for (auto& x : ev)
      if (x) (void)hipEventDestroy(x);
  };
  auto fail = [&](hipError_t e) {
    g_last_hip_error = static_cast<int>(e);
    release();
    return DBSCAN_HIP_ERR_HIP;
  };
  if (stage_ms) {
    for (auto& x : ev) {
      const hipError_t e = hipEventCreate(&x);
      if (e != hipSuccess) return fail(e);
    }
  }
  // Records the events due before kernel slot k (k == kNumKernels: the end).
  auto before = [&](int k) -> hipError_t {
    hipError_t e = hipSuccess;
    if (kev) e = hipEventRecord(kev[k], stream);
    if (stage_ms && e == hipSuccess) e = hipEventRecord(ev[k], stream);
    return e;
  };
  hipError_t e;

  if (ms < 1) {  // dbscan_grid.hpp:141: every point is noise
    for (int k = 0; k < kNumKernels; ++k)
      if ((e = before(k)) != hipSuccess) return fail(e);
    k_fill_i32<<<grid_for(n, 256, 4096), 256, 0, stream>>>(labels, n, kNoise, error_out);
    if ((e = before(kNumKernels)) != hipSuccess) return fail(e);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    if (stage_ms && (e = hipEventSynchronize(ev[kNumKernels])) != hipSuccess) return fail(e);
    release();
    return DBSCAN_HIP_OK;
  }

  void* ws = workspace;
  Params* prm = at<Params>(ws, L.params);
  const float2* X2 = reinterpret_cast<const float2*>(X);

  // A small cloud: one workgroup, one launch (the profile and per-kernel
  // timing variants report the pipeline's stages, so they keep it).
  if (n <= cfg.small_max && !stage_ms && !kev) {
    CloudInfo one;
    std::memset(&one, 0, sizeof(one));
    one.begin = 0;
    one.end = static_cast<int32_t>(n);
    one.inv_r = inv_r;
    one.eps_sq = eps_sq;
    one.best_init = best_init;
    one.ms = ms;
    k_small<<<1, kSmallThreads, 0, stream>>>(X2, labels, nullptr, 1, one, prm,
                                             take_injected_error(), error_out, n);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    release();
    return DBSCAN_HIP_OK;
  }

  uint64_t* sort_status = at<uint64_t>(ws, L.sort_status);
  float* xs_a = at<float>(ws, L.xs_a);
  float* ys_a = at<float>(ws, L.ys_a);
  uint32_t* idx_a = at<uint32_t>(ws, L.idx_a);
  float* xs_b = at<float>(ws, L.xs_b);
  float* ys_b = at<float>(ws, L.ys_b);
  uint32_t* idx_b = at<uint32_t>(ws, L.idx_b);
  int32_t* point_cell = at<int32_t>(ws, L.point_cell);
  uint8_t* is_core = at<uint8_t>(ws, L.is_core);
  int32_t* cell_start = at<int32_t>(ws, L.cell_start);
  uint64_t* cell_key = at<uint64_t>(ws, L.cell_key);
  uint32_t* cell_flags = at<uint32_t>(ws, L.cell_flags);
  uint4* cell_box = at<uint4>(ws, L.cell_box);
  int32_t* ngh = at<int32_t>(ws, L.ngh);
  int32_t* cell_ref = at<int32_t>(ws, L.cell_ref);
  int32_t* uf = at<int32_t>(ws, L.uf);
  uint32_t* rank = at<uint32_t>(ws, L.rank);
  int32_t* cell_cluster = at<int32_t>(ws, L.cell_cluster);

  const int g_reduce = grid_for(n, 2048, kMaxReduceBlocks);
  const int g_hist = grid_for(n, 16 * kHistThreads, kHistBlocks);
  const int g_sort = grid_for(n, kSortThreads * cfg.sort_items, INT_MAX);
  const int g_scan = grid_for(n, kScanTile, INT_MAX);
  const int g_celltiles = grid_for(n, kCellTile, INT_MAX);
  const int g_cells = grid_for(n, 256, 16384);
  const int g_cellwaves = grid_for(n, 256, cfg.cell_blocks);  // one lane per cell, grid-stride
  const int g_union = grid_for(n * 7, 256, cfg.pair_blocks);
  const int g_union_local = grid_for(n, kUnionRange, 4096);

  // Stage 1: cell keys (bounds + digit histograms).
  if ((e = before(kKBounds)) != hipSuccess) return fail(e);
  const int64_t zero16 = static_cast<int64_t>((L.zero_end - kParamsZeroBegin) / 16);
  k_bounds<<<g_reduce, 256, 0, stream>>>(X2, n, inv_r, at<int4>(ws, L.partials),
                                         at<uint4>(ws, kParamsZeroBegin), zero16);
  if ((e = before(kKHist)) != hipSuccess) return fail(e);
  k_hist<<<g_hist, kHistThreads, 0, stream>>>(X2, n, inv_r, at<int4>(ws, L.partials), g_reduce, prm,
                                     at<int2>(ws, L.col_range), L.max_cols, cfg.radix_bits,
                                              take_injected_error());

  // Stage 2: radix passes (sort + gather in one).
  if ((e = before(kKSort0)) != hipSuccess) return fail(e);
  launch_sort<true, false>(cfg.sort_items, g_sort, stream, X2, nullptr, nullptr, nullptr, xs_a,
                           ys_a, idx_a, prm, sort_status, 0, n, inv_r, nullptr, nullptr,
                           nullptr);
  const int tail_from = tail_first_pass(g_sort);
  for (int p = 1; p < kMaxPasses; ++p) {
    if ((e = before(kKSort0 + p)) != hipSuccess) return fail(e);
    const bool from_a = (p & 1) != 0;
    if (p < tail_from)
      launch_sort<false, false>(
          cfg.sort_items, g_sort, stream, nullptr, from_a ? xs_a : xs_b, from_a ? ys_a : ys_b,
          from_a ? idx_a : idx_b, from_a ? xs_b : xs_a, from_a ? ys_b : ys_a,
          from_a ? idx_b : idx_a, prm, sort_status, p, n, inv_r, nullptr, nullptr, nullptr);
    else if (p == tail_from)
      launch_sort_tail<false>(cfg.sort_items, g_sort, tail_from, stream, xs_a, ys_a, idx_a, xs_b,
                              ys_b, idx_b, prm, sort_status, n, inv_r, nullptr, nullptr, nullptr);
  }

  // Stage 3: cell runs.
  if ((e = before(kKCells)) != hipSuccess) return fail(e);
  k_cells<false><<<g_celltiles, kScanThreads, 0, stream>>>(
      xs_a, ys_a, xs_b, ys_b, prm, at<uint64_t>(ws, L.cell_status), point_cell, cell_start,
      cell_key, cell_flags, cell_box, at<int2>(ws, L.col_range), L.max_cols, n, inv_r, nullptr,
      nullptr, nullptr);
  const int g_boxes = grid_for(n, kBoxTile, INT_MAX);
  if ((e = before(kKBbox)) != hipSuccess) return fail(e);
  if (!kev && !stage_ms && n <= kFuseBoxNghMax) {
    // Stages 3 (cell boxes) and 4 (window columns) in one launch.
    k_bbox_ngh<false><<<g_boxes + g_cells, 256, 0, stream>>>(
        prm, xs_a, ys_a, xs_b, ys_b, point_cell, cell_start, cell_box, cell_flags, is_core, n, pp,
        static_cast<uint32_t>(g_boxes), cell_key, at<int2>(ws, L.col_range), L.max_cols, ngh,
        cell_ref, uf, nullptr, nullptr, xs_a, xs_b, static_cast<uint32_t>(g_cells));
    if ((e = before(kKNgh)) != hipSuccess) return fail(e);
  } else {
    k_bbox<<<g_boxes, 256, 0, stream>>>(prm, xs_a, ys_a, xs_b, ys_b, point_cell, cell_start,
                                        cell_box, cell_flags, is_core, n, pp);

    // Stage 4: window columns.
    if ((e = before(kKNgh)) != hipSuccess) return fail(e);
    k_ngh<false><<<g_cells, 256, 0, stream>>>(prm, cell_key, at<int2>(ws, L.col_range),
                                              L.max_cols, ngh, cell_ref, uf, nullptr, nullptr,
                                              xs_a, xs_b);
  }

  // Stage 5: core marking.
  if ((e = before(kKCore)) != hipSuccess) return fail(e);
  k_core_sparse<<<g_cellwaves, 256, 0, stream>>>(prm, xs_a, ys_a, xs_b, ys_b, cell_start, ngh,
                                                 cell_box, is_core, cell_flags, pp);

  // Stage 6: union-find.
  if ((e = before(kKUnion)) != hipSuccess) return fail(e);
  k_union<0><<<g_union, 256, 0, stream>>>(prm, xs_a, ys_a, xs_b, ys_b, cell_start, ngh, cell_ref,
                                          cell_flags, cell_box, cell_key, is_core, uf, pp, xs_a,
                                          xs_b);
  if ((e = before(kKUnionLocal)) != hipSuccess) return fail(e);
  k_union_local<<<g_union_local, kUnionLocalThreads, 0, stream>>>(prm, ngh, cell_ref, uf, xs_a,
                                                                  xs_b);
  if ((e = before(kKFlatten0)) != hipSuccess) return fail(e);
  k_flatten<<<g_cells, 256, 0, stream>>>(prm, cell_ref, cell_flags, uf, rank);
  if ((e = before(kKUnion1)) != hipSuccess) return fail(e);
  k_union<1><<<g_union, 256, 0, stream>>>(prm, xs_a, ys_a, xs_b, ys_b, cell_start, ngh, cell_ref,
                                          cell_flags, cell_box, cell_key, is_core, uf, pp, xs_a,
                                          xs_b);

  // Stage 7: cluster ids.
  if ((e = before(kKRank)) != hipSuccess) return fail(e);
  k_rank<<<g_scan, kScanThreads, 0, stream>>>(prm, at<uint64_t>(ws, L.rank_status), rank, uf);
  if ((e = before(kKCluster)) != hipSuccess) return fail(e);
  k_cluster<false><<<g_cells, 256, 0, stream>>>(prm, cell_ref, cell_flags, uf, rank,
                                                cell_cluster, nullptr, nullptr);

  // Stages 8-10: border assignment (sorted order), then core labels and the
  // unpermute in two bucketed passes.
  if ((e = before(kKBorder)) != hipSuccess) return fail(e);
  k_border<<<g_cellwaves, 256, 0, stream>>>(prm, xs_a, ys_a, idx_a, xs_b, ys_b, idx_b, cell_start,
                                            ngh, cell_flags, cell_box, is_core, cell_cluster, pp);
  if ((e = before(kKLabel)) != hipSuccess) return fail(e);
  const uint32_t bw = label_buckets(n);
  if (bw > 0) {
    launch_label(bw, n, stream, prm, idx_a, idx_b, point_cell, is_core, cell_cluster, xs_a, xs_b,
                 ys_a, ys_b, sort_status, at<uint32_t>(ws, L.long_runs));
  } else {
    k_label_direct<<<grid_for(n, 256, INT_MAX), 256, 0, stream>>>(
        prm, idx_a, idx_b, point_cell, is_core, cell_cluster, idx_a, idx_b, labels, n, error_out);
  }
  if ((e = before(kKUnpermute)) != hipSuccess) return fail(e);
  if (bw > 0)
    launch_unpermute(bw, n, stream, prm, xs_a, xs_b, ys_a, ys_b, sort_status,
                     at<uint32_t>(ws, L.long_runs), labels, error_out);
  if ((e = before(kNumKernels)) != hipSuccess) return fail(e);
  if ((e = hipGetLastError()) != hipSuccess) return fail(e);

  if (stage_ms) {
    if ((e = hipEventSynchronize(ev[kNumKernels])) != hipSuccess) return fail(e);
    for (int s = 0; s < DBSCAN_HIP_NUM_STAGES; ++s) {
      if ((e = hipEventElapsedTime(&stage_ms[s], ev[kStageSlots[s][0]], ev[kStageSlots[s][1]])) !=
          hipSuccess)
        return fail(e);
    }
  }
  release();
  return DBSCAN_HIP_OK;
}

// Per-cloud parameters as a single call computes them (see run()).
bool cloud_params(double eps, int64_t min_samples64, CloudInfo& ci) {
  if (!(eps > 0.0) || min_samples64 < 1) return false;
  const float eps_f = static_cast<float>(eps);
  const int32_t ms =
      static_cast<int32_t>(static_cast<uint32_t>(static_cast<uint64_t>(min_samples64)));
  ci.inv_r = static_cast<float>(std::sqrt(2.0)) / eps_f;
  ci.eps_sq = eps_f * eps_f;
  ci.best_init = ci.eps_sq + 1.0f;
  // a narrowed min_samples below 1 makes every point noise
  // (dbscan_grid.hpp:141); no cell of a batch reaches INT32_MAX points
  ci.ms = ms < 1 ? INT32_MAX : ms;
  return true;
}

// A batch's host-side tables: per-cloud records (validated exactly as a single
// call validates its arguments) and the tile counts the launches need.
struct BatchTables {
  std::vector<CloudInfo> info;
  dbscan_hip_batch_plan plan;
};

int build_batch(int64_t total, const int64_t* offsets, const double* eps,
                const int64_t* min_samples, int64_t num_clouds, int* bad_cloud,
                const dbscan_hip_config* cfgp, BatchTables& bt) {
  if (bad_cloud) *bad_cloud = -1;
  dbscan_hip_batch_plan& plan = bt.plan;
  std::memset(&plan, 0, sizeof(plan));
  const dbscan_hip_config cfg = cfgp ? *cfgp : default_config(total);
  if (!config_ok(cfg)) return DBSCAN_HIP_ERR_CONFIG;
  const int32_t sort_tile = kSortThreads * cfg.sort_items;
  if (num_clouds < 0 || num_clouds > (1 << 24)) return DBSCAN_HIP_ERR_N;
  if (total < 0 || total >= INT32_MAX) return DBSCAN_HIP_ERR_N;
  if (num_clouds == 0) return total == 0 ? DBSCAN_HIP_OK : DBSCAN_HIP_ERR_N;
  if (!offsets || !eps || !min_samples) return DBSCAN_HIP_ERR_NULL;
  const int32_t B = static_cast<int32_t>(num_clouds);
  bt.info.assign(B, CloudInfo{});
  int64_t sort_tiles = 0, scan_tiles = 0;
  if (offsets[0] != 0 || offsets[B] != total) return DBSCAN_HIP_ERR_N;
  for (int32_t i = 0; i < B; ++i) {
    CloudInfo& ci = bt.info[i];
    std::memset(&ci, 0, sizeof(ci));
    if (offsets[i + 1] < offsets[i]) {
      if (bad_cloud) *bad_cloud = i;
      return DBSCAN_HIP_ERR_N;
    }
    if (!cloud_params(eps[i], min_samples[i], ci)) {
      if (bad_cloud) *bad_cloud = i;
      return !(eps[i] > 0.0) ? DBSCAN_HIP_ERR_EPS : DBSCAN_HIP_ERR_MIN_SAMPLES;
    }
    ci.begin = static_cast<int32_t>(offsets[i]);
    ci.end = static_cast<int32_t>(offsets[i + 1]);
    ci.kxmin = INT_MAX; ci.kxmax = INT_MIN; ci.kymin = INT_MAX; ci.kymax = INT_MIN;
    ci.col_base = -1;
    const int64_t ni = offsets[i + 1] - offsets[i];
    sort_tiles += (ni + sort_tile - 1) / sort_tile;
    scan_tiles += (ni + kCellTile - 1) / kCellTile;
  }
  plan.total_points = total;
  plan.num_clouds = num_clouds;
  plan.sort_tiles = sort_tiles;
  plan.scan_tiles = scan_tiles;
  plan.cfg = cfg;
  return DBSCAN_HIP_OK;
}

// Copies the staged tables into the workspace on `stream` and decides the
// path: every cloud small, and at most kSmallGridClouds per CU, takes a
// persistent grid of one-workgroup clusterings that draws the clouds from a
// device ticket (allow_small false -- the per-kernel timing variant -- keeps
// the pipeline).  Past that many clouds per CU the pipeline catches up (on
// one MI355X, clouds of 200-4,096 points: 64 clouds 0.15 against 0.29 ms,
// 1,024 clouds 0.35 against 0.46, 2,048 clouds 0.62 against 0.63).  The
// host buffers are pageable, and a pageable host-to-device copy has read its
// source when it returns, so they may go out of scope at once.
int copy_batch_tables(BatchTables& bt, void* ws, hipStream_t stream, bool allow_small) {
  dbscan_hip_batch_plan& plan = bt.plan;
  const int32_t B = static_cast<int32_t>(plan.num_clouds);
  const dbscan_hip_config& cfg = plan.cfg;
  const Layout L = make_layout(plan.total_points, plan.num_clouds);
  const std::vector<CloudInfo>& info = bt.info;
  hipError_t e;
  bool all_small = allow_small && cfg.small_max > 0 && B <= kSmallGridClouds * cu_count();
  for (int32_t i = 0; i < B && all_small; ++i)
    all_small = info[i].end - info[i].begin <= cfg.small_max;
  plan.all_small = all_small ? 1 : 0;
  if (all_small) {
    // Largest clouds first (each cloud writes its labels at its own offset,
    // so the order is free): the workgroups that draw the last tickets then
    // get the shortest clouds, and the grid finishes together.
    std::vector<CloudInfo> order(info);
    std::stable_sort(order.begin(), order.end(), [](const CloudInfo& a, const CloudInfo& b) {
      return a.end - a.begin > b.end - b.begin;
    });
    e = hipMemcpyAsync(at<CloudInfo>(ws, L.clouds), order.data(), sizeof(CloudInfo) * B,
                       hipMemcpyHostToDevice, stream);
  } else {
    // Cloud records and tile tables, staged on the host and copied in one go
    // (the copy is ordered before the kernels on `stream`).
    const int32_t sort_tile = kSortThreads * cfg.sort_items;
    std::vector<char> staging(L.seg_end - L.clouds, 0);
    std::memcpy(staging.data(), info.data(), sizeof(CloudInfo) * B);
    SegTile* st_sort = reinterpret_cast<SegTile*>(staging.data() + (L.seg_sort - L.clouds));
    SegTile* st_scan = reinterpret_cast<SegTile*>(staging.data() + (L.seg_scan - L.clouds));
    int32_t a = 0, b = 0;
    for (int32_t i = 0; i < B; ++i) {
      const int32_t first = a;
      for (int32_t p0 = info[i].begin; p0 < info[i].end; p0 += sort_tile)
        st_sort[a++] = SegTile{i, p0, std::min(sort_tile, info[i].end - p0), first};
      for (int32_t p0 = info[i].begin; p0 < info[i].end; p0 += kCellTile)
        st_scan[b++] = SegTile{i, p0, std::min(kCellTile, info[i].end - p0), 0};
    }
    e = hipMemcpyAsync(at<char>(ws, L.clouds), staging.data(), staging.size(),
                       hipMemcpyHostToDevice, stream);
  }
  if (e != hipSuccess) {
    g_last_hip_error = static_cast<int>(e);
    return DBSCAN_HIP_ERR_HIP;
  }
  return DBSCAN_HIP_OK;
}

// The per-cloud cell bounds are accumulated with atomics by k_bounds_seg: a
// launch sequence that may be replayed (a captured graph) starts from the
// values the staging wrote.
__global__ __launch_bounds__(256) void k_clouds_reset(CloudInfo* __restrict__ clouds,
                                                      int32_t num_clouds) {
  const int32_t i = static_cast<int32_t>(blockIdx.x) * 256 + static_cast<int32_t>(threadIdx.x);
  if (i >= num_clouds) return;
  CloudInfo& ci = clouds[i];
  ci.kxmin = INT_MAX;
  ci.kxmax = INT_MIN;
  ci.kymin = INT_MAX;
  ci.kymax = INT_MIN;
}

// Kernels only (no copy, allocation or host synchronisation): the batch
// whose tables copy_batch_tables wrote into the workspace.  reset_bounds:
// the sequence may run again on the same tables (dbscan2d_hip_batched_enqueue,
// e.g. replays of a captured graph), so it first restores the per-cloud
// bounds the staging wrote.
int launch_batched(const float* X, int32_t* labels, void* workspace, hipStream_t stream,
                   const dbscan_hip_batch_plan& plan, hipEvent_t* kev, int32_t* error_out,
                   bool reset_bounds) {
  const int64_t total = plan.total_points;
  const int32_t B = static_cast<int32_t>(plan.num_clouds);
  const dbscan_hip_config& cfg = plan.cfg;
  const Layout L = make_layout(total, plan.num_clouds);
  auto fail = [&](hipError_t e) {
    g_last_hip_error = static_cast<int>(e);
    return DBSCAN_HIP_ERR_HIP;
  };
  hipError_t e;
  void* ws = workspace;
  const float2* X2 = reinterpret_cast<const float2*>(X);
  Params* prm = at<Params>(ws, L.params);
  if (plan.all_small) {
    CloudInfo* clouds = at<CloudInfo>(ws, L.clouds);
    const int grid = static_cast<int>(std::min<int64_t>(B, cu_count()));
    const int32_t error0 = take_injected_error();
    if (grid > 1) k_small_init<<<1, 64, 0, stream>>>(prm, error0);
    k_small<<<grid, kSmallThreads, 0, stream>>>(X2, labels, clouds, B, CloudInfo{}, prm,
                                                grid > 1 ? 0 : error0, error_out, total);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    return DBSCAN_HIP_OK;
  }
  auto before = [&](int k) -> hipError_t {
    return kev ? hipEventRecord(kev[k], stream) : hipSuccess;
  };
  if (reset_bounds)
    k_clouds_reset<<<(B + 255) / 256, 256, 0, stream>>>(at<CloudInfo>(ws, L.clouds), B);
  const int64_t sort_tiles = plan.sort_tiles, scan_tiles = plan.scan_tiles;
  const int64_t n = total;
  uint64_t* sort_status = at<uint64_t>(ws, L.sort_status);
  uint32_t* cloud_hist = at<uint32_t>(ws, L.cloud_hist);
  CloudInfo* clouds = at<CloudInfo>(ws, L.clouds);
  const SegTile* seg_sort = at<SegTile>(ws, L.seg_sort);
  const SegTile* seg_scan = at<SegTile>(ws, L.seg_scan);
  float* xs_a = at<float>(ws, L.xs_a);
  float* ys_a = at<float>(ws, L.ys_a);
  uint32_t* idx_a = at<uint32_t>(ws, L.idx_a);
  float* xs_b = at<float>(ws, L.xs_b);
  float* ys_b = at<float>(ws, L.ys_b);
  uint32_t* idx_b = at<uint32_t>(ws, L.idx_b);
  int32_t* point_cell = at<int32_t>(ws, L.point_cell);
  uint8_t* is_core = at<uint8_t>(ws, L.is_core);
  int32_t* cell_start = at<int32_t>(ws, L.cell_start);
  uint64_t* cell_key = at<uint64_t>(ws, L.cell_key);
  uint32_t* cell_flags = at<uint32_t>(ws, L.cell_flags);
  uint4* cell_box = at<uint4>(ws, L.cell_box);
  int32_t* ngh = at<int32_t>(ws, L.ngh);
  int32_t* cell_ref = at<int32_t>(ws, L.cell_ref);
  int32_t* uf = at<int32_t>(ws, L.uf);
  uint32_t* rank = at<uint32_t>(ws, L.rank);
  int32_t* cell_cluster = at<int32_t>(ws, L.cell_cluster);
  int32_t* cell_cloud = at<int32_t>(ws, L.cell_cloud);
  int2* col_range = at<int2>(ws, L.col_range);
  const ManyClouds pp{cell_cloud, clouds};

  const int g_sort = static_cast<int>(sort_tiles);
  const int g_scan_pts = static_cast<int>(scan_tiles);
  const int g_scan_cells = grid_for(n, kScanTile, INT_MAX);
  const int g_cells = grid_for(n, 256, 16384);
  const int g_cellwaves = grid_for(n, 256, cfg.cell_blocks);
  const int g_union = grid_for(n * 7, 256, cfg.pair_blocks);
  const int g_union_local = grid_for(n, kUnionRange, 4096);
  const int hist_per_block = static_cast<int>((sort_tiles + 1023) / 1024);
  const int g_hist = static_cast<int>((sort_tiles + hist_per_block - 1) / hist_per_block);

  if ((e = before(kKBounds)) != hipSuccess) return fail(e);
  const int64_t zero16 = static_cast<int64_t>((L.zero_end - kParamsZeroBegin) / 16);
  k_bounds_seg<<<g_sort, 256, 0, stream>>>(X2, seg_sort, clouds, at<uint4>(ws, kParamsZeroBegin),
                                           zero16);
  k_keys_seg<<<1, 1024, 0, stream>>>(prm, clouds, B, L.max_cols, cfg.radix_bits,
                                     take_injected_error());
  if ((e = before(kKHist)) != hipSuccess) return fail(e);
  k_hist_seg<<<g_hist, 256, 0, stream>>>(X2, seg_sort, g_sort, hist_per_block, clouds, prm,
                                         cloud_hist, col_range);
  if ((e = before(kKSort0)) != hipSuccess) return fail(e);
  launch_sort<true, true>(cfg.sort_items, g_sort, stream, X2, nullptr, nullptr, nullptr, xs_a,
                          ys_a, idx_a, prm, sort_status, 0, n, 0.f, seg_sort, clouds,
                          cloud_hist);
  const int tail_from = tail_first_pass(g_sort);
  for (int p = 1; p < kMaxPasses; ++p) {
    if ((e = before(kKSort0 + p)) != hipSuccess) return fail(e);
    const bool from_a = (p & 1) != 0;
    if (p < tail_from)
      launch_sort<false, true>(
          cfg.sort_items, g_sort, stream, nullptr, from_a ? xs_a : xs_b, from_a ? ys_a : ys_b,
          from_a ? idx_a : idx_b, from_a ? xs_b : xs_a, from_a ? ys_b : ys_a,
          from_a ? idx_b : idx_a, prm, sort_status, p, n, 0.f, seg_sort, clouds, cloud_hist);
    else if (p == tail_from)
      launch_sort_tail<true>(cfg.sort_items, g_sort, tail_from, stream, xs_a, ys_a, idx_a, xs_b,
                             ys_b, idx_b, prm, sort_status, n, 0.f, seg_sort, clouds, cloud_hist);
  }
  if ((e = before(kKCells)) != hipSuccess) return fail(e);
  k_cells<true><<<g_scan_pts, kScanThreads, 0, stream>>>(
      xs_a, ys_a, xs_b, ys_b, prm, at<uint64_t>(ws, L.cell_status), point_cell, cell_start,
      cell_key, cell_flags, cell_box, col_range, L.max_cols, n, 0.f, seg_scan, clouds,
      cell_cloud);
  const int g_boxes = grid_for(n, kBoxTile, INT_MAX);
  if ((e = before(kKBbox)) != hipSuccess) return fail(e);
  if (!kev && n <= kFuseBoxNghMax) {
    k_bbox_ngh<true><<<g_boxes + g_cells, 256, 0, stream>>>(
        prm, xs_a, ys_a, xs_b, ys_b, point_cell, cell_start, cell_box, cell_flags, is_core, n, pp,
        static_cast<uint32_t>(g_boxes), cell_key, col_range, L.max_cols, ngh, cell_ref, uf,
        cell_cloud, clouds, xs_a, xs_b, static_cast<uint32_t>(g_cells));
    if ((e = before(kKNgh)) != hipSuccess) return fail(e);
  } else {
    k_bbox<<<g_boxes, 256, 0, stream>>>(prm, xs_a, ys_a, xs_b, ys_b, point_cell, cell_start,
                                        cell_box, cell_flags, is_core, n, pp);
    if ((e = before(kKNgh)) != hipSuccess) return fail(e);
    k_ngh<true><<<g_cells, 256, 0, stream>>>(prm, cell_key, col_range, L.max_cols, ngh,
                                             cell_ref, uf, cell_cloud, clouds, xs_a, xs_b);
  }
  if ((e = before(kKCore)) != hipSuccess) return fail(e);
  k_core_sparse<<<g_cellwaves, 256, 0, stream>>>(prm, xs_a, ys_a, xs_b, ys_b, cell_start, ngh,
                                                 cell_box, is_core, cell_flags, pp);
  if ((e = before(kKUnion)) != hipSuccess) return fail(e);
  k_union<0><<<g_union, 256, 0, stream>>>(prm, xs_a, ys_a, xs_b, ys_b, cell_start, ngh, cell_ref,
                                          cell_flags, cell_box, cell_key, is_core, uf, pp, xs_a,
                                          xs_b);
  if ((e = before(kKUnionLocal)) != hipSuccess) return fail(e);
  k_union_local<<<g_union_local, kUnionLocalThreads, 0, stream>>>(prm, ngh, cell_ref, uf, xs_a,
                                                                  xs_b);
  if ((e = before(kKFlatten0)) != hipSuccess) return fail(e);
  k_flatten<<<g_cells, 256, 0, stream>>>(prm, cell_ref, cell_flags, uf, rank);
  if ((e = before(kKUnion1)) != hipSuccess) return fail(e);
  k_union<1><<<g_union, 256, 0, stream>>>(prm, xs_a, ys_a, xs_b, ys_b, cell_start, ngh, cell_ref,
                                          cell_flags, cell_box, cell_key, is_core, uf, pp, xs_a,
                                          xs_b);
  if ((e = before(kKRank)) != hipSuccess) return fail(e);
  k_rank<<<g_scan_cells, kScanThreads, 0, stream>>>(prm, at<uint64_t>(ws, L.rank_status), rank,
                                                    uf);
  if ((e = before(kKCluster)) != hipSuccess) return fail(e);
  k_cluster<true><<<g_cells, 256, 0, stream>>>(prm, cell_ref, cell_flags, uf, rank, cell_cluster,
                                               cell_cloud, clouds);
  if ((e = before(kKBorder)) != hipSuccess) return fail(e);
  k_border<<<g_cellwaves, 256, 0, stream>>>(prm, xs_a, ys_a, idx_a, xs_b, ys_b, idx_b, cell_start,
                                            ngh, cell_flags, cell_box, is_core, cell_cluster, pp);
  if ((e = before(kKLabel)) != hipSuccess) return fail(e);
  const uint32_t bw = label_buckets(n);
  if (bw > 0) {
    launch_label(bw, n, stream, prm, idx_a, idx_b, point_cell, is_core, cell_cluster, xs_a, xs_b,
                 ys_a, ys_b, sort_status, at<uint32_t>(ws, L.long_runs));
  } else {
    k_label_direct<<<grid_for(n, 256, INT_MAX), 256, 0, stream>>>(
        prm, idx_a, idx_b, point_cell, is_core, cell_cluster, idx_a, idx_b, labels, n, error_out);
  }
  if ((e = before(kKUnpermute)) != hipSuccess) return fail(e);
  if (bw > 0)
    launch_unpermute(bw, n, stream, prm, xs_a, xs_b, ys_a, ys_b, sort_status,
                     at<uint32_t>(ws, L.long_runs), labels, error_out);
  if ((e = before(kNumKernels)) != hipSuccess) return fail(e);
  if ((e = hipGetLastError()) != hipSuccess) return fail(e);
  return DBSCAN_HIP_OK;
}
int run_batched(const float* X, int64_t total, const int64_t* offsets, const double* eps,
                const int64_t* min_samples, int64_t num_clouds, int32_t* labels,
                void* workspace, size_t workspace_bytes, hipStream_t stream, hipEvent_t* kev,
                int* bad_cloud, const dbscan_hip_config* cfgp, int32_t* error_out = nullptr) {
  BatchTables bt;
  const int rc = build_batch(total, offsets, eps, min_samples, num_clouds, bad_cloud, cfgp, bt);
  if (rc != DBSCAN_HIP_OK || total == 0) return rc;
  if (!X || !labels || !workspace) return DBSCAN_HIP_ERR_NULL;
  if ((reinterpret_cast<uintptr_t>(X) & 7u) || (reinterpret_cast<uintptr_t>(labels) & 3u) ||
      (reinterpret_cast<uintptr_t>(workspace) & 255u))
    return DBSCAN_HIP_ERR_ALIGN;
  if (workspace_bytes < make_layout(total, num_clouds).total) return DBSCAN_HIP_ERR_WORKSPACE;
  const int rc2 = copy_batch_tables(bt, workspace, stream, kev == nullptr);
  if (rc2 != DBSCAN_HIP_OK) return rc2;
  return launch_batched(X, labels, workspace, stream, bt.plan, kev, error_out, false);
}

std::atomic<void*> g_scratch_addr[kMaxDevices];

// The current device's scratch address, looked up on its first use.
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    g_last_hip_error = static_cast<int>(e);
    return DBSCAN_HIP_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) return DBSCAN_HIP_ERR_N;
  void* p = g_scratch_addr[dev].load(std::memory_order_acquire);
  if (!p) {
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_last_hip_error = static_cast<int>(e != hipSuccess ? e : hipErrorInvalidSymbol);
      return DBSCAN_HIP_ERR_HIP;
    }
    g_scratch_addr[dev].store(p, std::memory_order_release);
  }
  *out = p;
  return DBSCAN_HIP_OK;
}

}  // namespace

extern "C" {

int dbscan2d_hip_workspace_bytes(int64_t n, size_t* bytes) {
  if (!bytes) return DBSCAN_HIP_ERR_NULL;
  if (n < 0 || n > INT32_MAX) return DBSCAN_HIP_ERR_N;
  *bytes = make_layout(n).total;
  return DBSCAN_HIP_OK;
}

int dbscan2d_hip(const float* X, int64_t n, double eps, int64_t min_samples, int32_t* labels,
                 void* workspace, size_t workspace_bytes, hipStream_t stream) {
  return run(X, n, eps, min_samples, labels, workspace, workspace_bytes, stream, nullptr,
             nullptr, nullptr);
}

int dbscan2d_hip_noalloc(const float* X, int64_t n, double eps, int64_t min_samples,
                         int32_t* labels, hipStream_t stream) {
  if (n < 0 || n > INT32_MAX) return DBSCAN_HIP_ERR_N;
  if (!(eps > 0.0)) return DBSCAN_HIP_ERR_EPS;
  if (min_samples < 1) return DBSCAN_HIP_ERR_MIN_SAMPLES;
  if (n == 0) return DBSCAN_HIP_OK;
  if (n > DBSCAN_HIP_NOALLOC_MAX_POINTS || make_layout(n).total > DBSCAN_HIP_NOALLOC_SCRATCH_BYTES)
    return DBSCAN_HIP_ERR_TOO_LARGE;
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc != DBSCAN_HIP_OK) return rc;
  return run(X, n, eps, min_samples, labels, ws, DBSCAN_HIP_NOALLOC_SCRATCH_BYTES, stream,
             nullptr, nullptr, nullptr);
}

int dbscan2d_hip_default_config(int64_t n, dbscan_hip_config* cfg) {
  if (!cfg) return DBSCAN_HIP_ERR_NULL;
  *cfg = default_config(n);
  return DBSCAN_HIP_OK;
}

int dbscan2d_hip_configured(const float* X, int64_t n, double eps, int64_t min_samples,
                            int32_t* labels, void* workspace, size_t workspace_bytes,
                            hipStream_t stream, const dbscan_hip_config* cfg) {
  return run(X, n, eps, min_samples, labels, workspace, workspace_bytes, stream, nullptr,
             nullptr, cfg);
}

int dbscan2d_hip_checked(const float* X, int64_t n, double eps, int64_t min_samples,
                         int32_t* labels, void* workspace, size_t workspace_bytes,
                         hipStream_t stream, const dbscan_hip_config* cfg, int32_t* error_out) {
  return run(X, n, eps, min_samples, labels, workspace, workspace_bytes, stream, nullptr,
             nullptr, cfg, error_out);
}

int dbscan2d_hip_timed(const float* X, int64_t n, double eps, int64_t min_samples,
                       int32_t* labels, void* workspace, size_t workspace_bytes,
                       hipStream_t stream, hipEvent_t* events, const dbscan_hip_config* cfg) {
  if (!events) return DBSCAN_HIP_ERR_NULL;
  return run(X, n, eps, min_samples, labels, workspace, workspace_bytes, stream, nullptr,
             events, cfg);
}

const char* dbscan2d_hip_kernel_name(int i) {
  return (i >= 0 && i < kNumKernels) ? kKernelNames[i] : nullptr;
}

int dbscan2d_hip_profile(const float* X, int64_t n, double eps, int64_t min_samples,
                         int32_t* labels, void* workspace, size_t workspace_bytes,
                         hipStream_t stream, float* stage_ms) {
  if (!stage_ms) return DBSCAN_HIP_ERR_NULL;
  return run(X, n, eps, min_samples, labels, workspace, workspace_bytes, stream, stage_ms,
             nullptr, nullptr);
}

const char* dbscan2d_hip_stage_name(int i) {
  return (i >= 0 && i < DBSCAN_HIP_NUM_STAGES) ? kStageNames[i] : nullptr;
}

int dbscan2d_hip_batched_workspace_bytes(int64_t total_points, int64_t num_clouds,
                                         size_t* bytes) {
  if (!bytes) return DBSCAN_HIP_ERR_NULL;
  if (total_points < 0 || total_points >= INT32_MAX || num_clouds < 0 ||
      num_clouds > (1 << 24))
    return DBSCAN_HIP_ERR_N;
  *bytes = make_layout(total_points, num_clouds > 0 ? num_clouds : 1).total;
  return DBSCAN_HIP_OK;
}

int dbscan2d_hip_batched(const float* X, int64_t total_points, const int64_t* offsets,
                         const double* eps, const int64_t* min_samples, int64_t num_clouds,
                         int32_t* labels, void* workspace, size_t workspace_bytes,
                         hipStream_t stream, int* bad_cloud, const dbscan_hip_config* cfg) {
  return run_batched(X, total_points, offsets, eps, min_samples, num_clouds, labels, workspace,
                     workspace_bytes, stream, nullptr, bad_cloud, cfg);
}

int dbscan2d_hip_batched_checked(const float* X, int64_t total_points, const int64_t* offsets,
                                 const double* eps, const int64_t* min_samples,
                                 int64_t num_clouds, int32_t* labels, void* workspace,
                                 size_t workspace_bytes, hipStream_t stream, int* bad_cloud,
                                 const dbscan_hip_config* cfg, int32_t* error_out) {
  return run_batched(X, total_points, offsets, eps, min_samples, num_clouds, labels, workspace,
                     workspace_bytes, stream, nullptr, bad_cloud, cfg, error_out);
}

int dbscan2d_hip_batched_timed(const float* X, int64_t total_points, const int64_t* offsets,
                               const double* eps, const int64_t* min_samples,
                               int64_t num_clouds, int32_t* labels, void* workspace,
                               size_t workspace_bytes, hipStream_t stream, hipEvent_t* events,
                               const dbscan_hip_config* cfg) {
  if (!events) return DBSCAN_HIP_ERR_NULL;
  return run_batched(X, total_points, offsets, eps, min_samples, num_clouds, labels, workspace,
                     workspace_bytes, stream, events, nullptr, cfg);
}

int dbscan2d_hip_batched_stage(int64_t total_points, const int64_t* offsets, const double* eps,
                               const int64_t* min_samples, int64_t num_clouds, void* workspace,
                               size_t workspace_bytes, hipStream_t stream, int* bad_cloud,
                               const dbscan_hip_config* cfg, dbscan_hip_batch_plan* plan) {
  if (!plan) return DBSCAN_HIP_ERR_NULL;
  BatchTables bt;
  const int rc = build_batch(total_points, offsets, eps, min_samples, num_clouds, bad_cloud, cfg,
                             bt);
  *plan = bt.plan;
  if (rc != DBSCAN_HIP_OK || total_points == 0) return rc;
  if (!workspace) return DBSCAN_HIP_ERR_NULL;
  if (reinterpret_cast<uintptr_t>(workspace) & 255u) return DBSCAN_HIP_ERR_ALIGN;
  if (workspace_bytes < make_layout(total_points, num_clouds).total)
    return DBSCAN_HIP_ERR_WORKSPACE;
  const int rc2 = copy_batch_tables(bt, workspace, stream, true);
  *plan = bt.plan;
  return rc2;
}

int dbscan2d_hip_batched_enqueue(const float* X, int32_t* labels, void* workspace,
                                 size_t workspace_bytes, hipStream_t stream,
                                 const dbscan_hip_batch_plan* plan, int32_t* error_out) {
  if (!plan) return DBSCAN_HIP_ERR_NULL;
  const dbscan_hip_batch_plan& pl = *plan;
  if (pl.total_points < 0 || pl.total_points >= INT32_MAX || pl.num_clouds < 0 ||
      pl.num_clouds > (1 << 24) || !config_ok(pl.cfg))
    return DBSCAN_HIP_ERR_CONFIG;
  if (pl.total_points == 0) return DBSCAN_HIP_OK;
  if (!X || !labels || !workspace) return DBSCAN_HIP_ERR_NULL;
  if ((reinterpret_cast<uintptr_t>(X) & 7u) || (reinterpret_cast<uintptr_t>(labels) & 3u) ||
      (reinterpret_cast<uintptr_t>(workspace) & 255u))
    return DBSCAN_HIP_ERR_ALIGN;
  if (workspace_bytes < make_layout(pl.total_points, pl.num_clouds).total)
    return DBSCAN_HIP_ERR_WORKSPACE;
  return launch_batched(X, labels, workspace, stream, pl, nullptr, error_out, true);
}

int dbscan2d_hip_layout(int64_t n, dbscan_hip_layout* out) {
  if (!out) return DBSCAN_HIP_ERR_NULL;
  if (n < 0 || n > INT32_MAX) return DBSCAN_HIP_ERR_N;
  const Layout L = make_layout(n);
  out->info = L.params;
  out->xs_a = L.xs_a; out->ys_a = L.ys_a; out->idx_a = L.idx_a;
  out->xs_b = L.xs_b; out->ys_b = L.ys_b; out->idx_b = L.idx_b;
  out->point_cell = L.point_cell;
  out->is_core = L.is_core;
  out->cell_start = L.cell_start;
  out->cell_key = L.cell_key;
  out->cell_flags = L.cell_flags;
  out->cell_ref = L.cell_ref;
  out->uf = L.uf;
  out->cell_cluster = L.cell_cluster;
  return DBSCAN_HIP_OK;
}

const char* dbscan2d_hip_status_string(int status) {
  switch (status) {
    case DBSCAN_HIP_OK: return "ok";
    case DBSCAN_HIP_ERR_N: return "n must be in [0, 2^31)";
    case DBSCAN_HIP_ERR_EPS: return "eps must be > 0";
    case DBSCAN_HIP_ERR_MIN_SAMPLES: return "min_samples must be >= 1";
    case DBSCAN_HIP_ERR_NULL: return "null pointer argument";
    case DBSCAN_HIP_ERR_WORKSPACE: return "workspace too small";
    case DBSCAN_HIP_ERR_ALIGN: return "misaligned pointer (X 8 B, labels 4 B, workspace 256 B)";
    case DBSCAN_HIP_ERR_HIP: return "HIP runtime error";
    case DBSCAN_HIP_ERR_CONFIG: return "invalid launch configuration";
    case DBSCAN_HIP_ERR_DEVICE:
      return "a device-side check failed; every label of the call is INT32_MIN";
    case DBSCAN_HIP_ERR_TOO_LARGE:
      return "n exceeds DBSCAN_HIP_NOALLOC_MAX_POINTS, the limit of the allocation-free entry "
             "point (use dbscan2d_hip with a workspace)";
    default: return "unknown status";
  }
}

int dbscan2d_hip_last_hip_error(void) { return g_last_hip_error; }

int dbscan2d_hip_error_async(const void* workspace, int32_t* dst, hipStream_t stream) {
  if (!workspace || !dst) return DBSCAN_HIP_ERR_NULL;
  if (reinterpret_cast<uintptr_t>(workspace) & 255u) return DBSCAN_HIP_ERR_ALIGN;
  k_copy_error<<<1, 64, 0, stream>>>(static_cast<const Params*>(workspace), dst);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_last_hip_error = static_cast<int>(e);
    return DBSCAN_HIP_ERR_HIP;
  }
  return DBSCAN_HIP_OK;
}

int dbscan2d_hip_check(const void* workspace, hipStream_t stream, int32_t* bits) {
  if (bits) *bits = 0;
  if (!workspace) return DBSCAN_HIP_ERR_NULL;
  if (reinterpret_cast<uintptr_t>(workspace) & 255u) return DBSCAN_HIP_ERR_ALIGN;
  int32_t word = 0;
  hipError_t e = hipStreamSynchronize(stream);
  if (e == hipSuccess)
    e = hipMemcpy(&word, static_cast<const char*>(workspace) + offsetof(Params, error),
                  sizeof(word), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    g_last_hip_error = static_cast<int>(e);
    return DBSCAN_HIP_ERR_HIP;
  }
  if (bits) *bits = word;
  return word ? DBSCAN_HIP_ERR_DEVICE : DBSCAN_HIP_OK;
}

const char* dbscan2d_hip_device_error_name(int32_t bit) {
  switch (bit) {
    case DBSCAN_HIP_DEV_SORT_WAIT: return "a radix-sort look-back or pass wait timed out";
    case DBSCAN_HIP_DEV_SCAN_WAIT: return "a scan look-back wait timed out";
    case DBSCAN_HIP_DEV_UF_LOOP: return "a union-find operation reached its step limit";
    case DBSCAN_HIP_DEV_RANGE: return "a computed index fell outside its array";
    default: return nullptr;
  }
}

void dbscan2d_hip_inject_error_for_testing(int32_t bits) { g_injected_error = bits; }

}  // extern "C"
