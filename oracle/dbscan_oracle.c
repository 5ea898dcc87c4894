/*
 * dbscan_oracle.c -- CPU restatement of the reference's grid-hashed 2-D DBSCAN.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
 * path: it is linked by the test suite, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, and by nothing else.  The product library
 * (dbscan-torch_amd/) never links, loads or calls it.
 *
 * It restates, rule for rule, the semantic authority named by the reference:
 *   csrc/cpu/dbscan_grid.hpp   (the 10-stage pipeline, lines 136-458)
 *   csrc/cpu/dbscan_cpu.cpp    (casts, -1 fill, n == 0; lines 34-59)
 * in plain C11 without ATen.  Each stage cites the reference lines it follows.
 * Two mechanical substitutions, neither of which can change a label:
 *   - at::sort on the signed int64 key (dbscan_grid.hpp:176-178) becomes a
 *     stable LSD radix sort of the key with its sign bit flipped, which yields
 *     the same signed-ascending key order;
 *   - the unordered_map cell lookup (dbscan_grid.hpp:229-253) becomes a binary
 *     search over the sorted unique keys, which returns the same cell index or
 *     -1.
 * Within-cell point order may differ from at::sort's (which need not be
 * stable); no label depends on it (every core point of a cell carries the
 * cell's cluster id, and the border rule picks a label, not a point).
 *
 * The union-find runs serially.  The reference's lock-free version hooks the
 * larger root onto the smaller (dbscan_grid.hpp:106-116), so in both the root
 * of a component is its minimum cell index and the result is independent of
 * schedule.
 *
 * Float semantics: build with -ffp-contract=off.  x86-64 without -march has no
 * FMA, so the reference evaluates ex*ex + ey*ey as two rounded multiplies and
 * one rounded add; so does this file.  Float-to-int32 conversion of an
 * out-of-range or NaN value is undefined in C++; the reference's x86 build
 * (cvttss2si) yields INT32_MIN, and cell_coord() states that choice explicitly.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_NOISE (-1)

/* dbscan_grid.hpp:162-163 -- (int32)floor(x * inv_r), x86 conversion rule. */
static inline int32_t cell_coord(float v) {
  const float f = floorf(v);
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
  return (int32_t)f;
}

/* dbscan_grid.hpp:42-45 -- (u64)(u32)kx << 32 | (u32)ky. */
static inline uint64_t cell_key(int32_t kx, int32_t ky) {
  return ((uint64_t)(uint32_t)kx << 32) | (uint64_t)(uint32_t)ky;
}

/* std::min / std::max / std::max({a,b,c}) as the reference's libstdc++ calls
 * evaluate them (dbscan_grid.hpp:124-133, 338-341), NaN behaviour included. */
static inline float std_min(float a, float b) { return (b < a) ? b : a; }
static inline float std_max(float a, float b) { return (a < b) ? b : a; }
static inline float std_max3(float a, float b, float c) {
  float m = a;
  if (m < b) m = b;
  if (m < c) m = c;
  return m;
}

/* dbscan_grid.hpp:123-127 */
static inline float point_box_d2(float x, float y, const float* box) {
  const float dx = std_max3(box[0] - x, x - box[1], 0.0f);
  const float dy = std_max3(box[2] - y, y - box[3], 0.0f);
  return dx * dx + dy * dy;
}

/* dbscan_grid.hpp:130-134 */
static inline float box_box_d2(const float* a, const float* b) {
  const float dx = std_max3(b[0] - a[1], a[0] - b[1], 0.0f);
  const float dy = std_max3(b[2] - a[3], a[2] - b[3], 0.0f);
  return dx * dx + dy * dy;
}

/* Serial union-find with the reference's hooking rule (dbscan_grid.hpp:95-116):
 * path-halving find, larger root hooked onto the smaller. */
static int32_t uf_find(int32_t* uf, int32_t x) {
  int32_t p = uf[x];
  while (p != x) {
    const int32_t gp = uf[p];
    uf[x] = gp;
    x = p;
    p = gp;
  }
  return x;
}

static void uf_unite(int32_t* uf, int32_t a, int32_t b) {
  a = uf_find(uf, a);
  b = uf_find(uf, b);
  if (a == b) return;
  if (a < b) uf[b] = a;
  else uf[a] = b;
}

/* Stable argsort of signed int64 keys, ascending (the order at::sort(kLong)
 * produces at dbscan_grid.hpp:176-178).  LSD radix over 16-bit digits of the
 * sign-flipped key; passes whose digit is constant are skipped. */
static int argsort_i64(const uint64_t* keys, int32_t n, int32_t* perm) {
  uint64_t* ka = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
  uint64_t* kb = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
  int32_t* pb = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  size_t* cnt = (size_t*)malloc(sizeof(size_t) * 65536);
  if (!ka || !kb || !pb || !cnt) {
    free(ka); free(kb); free(pb); free(cnt);
    return 1;
  }
  for (int32_t i = 0; i < n; ++i) {
    ka[i] = keys[i] ^ 0x8000000000000000ull;
    perm[i] = i;
  }
  uint64_t* src_k = ka; uint64_t* dst_k = kb;
  int32_t* src_p = perm; int32_t* dst_p = pb;
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 16 * pass;
    memset(cnt, 0, sizeof(size_t) * 65536);
    for (int32_t i = 0; i < n; ++i) cnt[(src_k[i] >> shift) & 0xFFFFu]++;
    if (cnt[(src_k[0] >> shift) & 0xFFFFu] == (size_t)n) continue;
    size_t sum = 0;
    for (int d = 0; d < 65536; ++d) { size_t c = cnt[d]; cnt[d] = sum; sum += c; }
    for (int32_t i = 0; i < n; ++i) {
      const size_t pos = cnt[(src_k[i] >> shift) & 0xFFFFu]++;
      dst_k[pos] = src_k[i];
      dst_p[pos] = src_p[i];
    }
    uint64_t* tk = src_k; src_k = dst_k; dst_k = tk;
    int32_t* tp = src_p; src_p = dst_p; dst_p = tp;
  }
  if (src_p != perm) memcpy(perm, src_p, sizeof(int32_t) * (size_t)n);
  free(ka); free(kb); free(pb); free(cnt);
  return 0;
}

/* Lower bound over the sorted unique keys, compared as signed int64. */
static int32_t find_cell(const int64_t* ukeys, int32_t nc, int64_t target) {
  int32_t lo = 0, hi = nc;
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (ukeys[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  return (lo < nc && ukeys[lo] == target) ? lo : -1;
}

typedef struct {
  int32_t* perm;           /* [n]    sorted position -> original index      */
  uint8_t* is_core;        /* [n]    core flag, sorted order                */
  int32_t* cell_start;     /* [n+1]  first num_cells+1 entries used         */
  int32_t* cell_kx;        /* [n]    cell coordinates, reference cell order */
  int32_t* cell_ky;        /* [n]                                           */
  uint8_t* cell_status;    /* [n]    0 no core, 1 mixed, 2 all core          */
  int32_t* cell_root;      /* [n]    union-find root (min cell index)        */
  int32_t* cell_cluster;   /* [n]    cluster id of the cell, -1 if no core   */
  int32_t num_cells;
  int32_t n_clusters;
} oracle_debug_t;

/*
 * Core pipeline.  Returns 0 on success, 1 on allocation failure, 2 on bad
 * arguments.  `labels` must hold n entries; it is filled with -1 first
 * (dbscan_cpu.cpp:42) and left so when n == 0 or min_samples < 1
 * (dbscan_grid.hpp:141).
 */
static int run(const float* X, int64_t n64, double eps, int64_t min_samples64,
               int32_t* labels, oracle_debug_t* dbg) {
  if (n64 < 0 || n64 > INT32_MAX || (n64 > 0 && (!X || !labels))) return 2;
  const int32_t n = (int32_t)n64;                       /* dbscan_cpu.cpp:41 */
  for (int32_t i = 0; i < n; ++i) labels[i] = ORACLE_NOISE;
  if (dbg) { dbg->num_cells = 0; dbg->n_clusters = 0; }
  const float eps_f = (float)eps;                        /* dbscan_cpu.cpp:52 */
  const int32_t min_samples = (int32_t)(uint32_t)(uint64_t)min_samples64;  /* :53 */
  if (n == 0 || min_samples < 1) return 0;               /* dbscan_grid.hpp:141 */

  const float inv_r = (float)sqrt(2.0) / eps_f;          /* dbscan_grid.hpp:153 */
  const float eps_sq = eps_f * eps_f;                    /* dbscan_grid.hpp:154 */

  int rc = 1;
  uint64_t* keys = NULL; int32_t* perm = NULL; int64_t* skeys = NULL;
  float *xs = NULL, *ys = NULL, *bbox = NULL;
  int32_t *cell_start = NULL, *cell_kx = NULL, *cell_ky = NULL, *ngh = NULL;
  int32_t *point_cell = NULL, *uf = NULL, *root_to_cluster = NULL, *sorted_labels = NULL;
  uint8_t *is_core = NULL, *cell_status = NULL;

  /* split_xy (dbscan_cpu.cpp:17-25) is folded into the reads of X below. */

  /* Stage 1: per-point packed cell key (dbscan_grid.hpp:159-166). */
  keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
  if (!keys) goto done;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t kx = cell_coord(X[2 * (size_t)i] * inv_r);
    const int32_t ky = cell_coord(X[2 * (size_t)i + 1] * inv_r);
    keys[i] = cell_key(kx, ky);
  }

  /* Stage 2: sort by signed key, gather xs/ys (dbscan_grid.hpp:173-201). */
  perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  skeys = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  xs = (float*)malloc(sizeof(float) * (size_t)n);
  ys = (float*)malloc(sizeof(float) * (size_t)n);
  if (!perm || !skeys || !xs || !ys) goto done;
  if (argsort_i64(keys, n, perm)) goto done;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t p = perm[i];
    skeys[i] = (int64_t)keys[p];
    xs[i] = X[2 * (size_t)p];
    ys[i] = X[2 * (size_t)p + 1];
  }
  free(keys); keys = NULL;

  /* Stage 3: cell boundaries by a sequential walk (dbscan_grid.hpp:206-227). */
  cell_start = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  cell_kx = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  cell_ky = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int64_t* ukeys = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  if (!cell_start || !cell_kx || !cell_ky || !ukeys) { free(ukeys); goto done; }
  int32_t num_cells = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (i == 0 || skeys[i] != skeys[i - 1]) {
      const uint64_t k = (uint64_t)skeys[i];
      cell_start[num_cells] = i;
      cell_kx[num_cells] = (int32_t)(uint32_t)(k >> 32);
      cell_ky[num_cells] = (int32_t)(uint32_t)k;
      ukeys[num_cells] = skeys[i];
      ++num_cells;
    }
  }
  cell_start[num_cells] = n;

  /* Stage 4: 5x5 neighbour table, off = (dx+2)*5 + (dy+2), dx-major
   * (dbscan_grid.hpp:239-252).  kx+dx wraps as 32-bit arithmetic does. */
  ngh = (int32_t*)malloc(sizeof(int32_t) * (size_t)num_cells * 25);
  if (!ngh) { free(ukeys); goto done; }
#pragma omp parallel for schedule(static)
  for (int32_t c = 0; c < num_cells; ++c) {
    for (int32_t dx = -2; dx <= 2; ++dx) {
      for (int32_t dy = -2; dy <= 2; ++dy) {
        const int32_t nx = (int32_t)((uint32_t)cell_kx[c] + (uint32_t)dx);
        const int32_t ny = (int32_t)((uint32_t)cell_ky[c] + (uint32_t)dy);
        ngh[(size_t)c * 25 + (size_t)((dx + 2) * 5 + (dy + 2))] =
            find_cell(ukeys, num_cells, (int64_t)cell_key(nx, ny));
      }
    }
  }
  free(ukeys);

  /* Stage 5: point -> cell, core marking passes 1 and 2, cell status
   * (dbscan_grid.hpp:258-321). */
  point_cell = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  is_core = (uint8_t*)calloc((size_t)n, 1);
  cell_status = (uint8_t*)calloc((size_t)num_cells, 1);
  if (!point_cell || !is_core || !cell_status) goto done;
  for (int32_t c = 0; c < num_cells; ++c)
    for (int32_t i = cell_start[c]; i < cell_start[c + 1]; ++i) point_cell[i] = c;
  for (int32_t c = 0; c < num_cells; ++c)                       /* pass 1, :271-279 */
    if (cell_start[c + 1] - cell_start[c] >= min_samples)
      for (int32_t i = cell_start[c]; i < cell_start[c + 1]; ++i) is_core[i] = 1;
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t i = 0; i < n; ++i) {                             /* pass 2, :284-307 */
    if (is_core[i]) continue;
    const int32_t* row = ngh + (size_t)point_cell[i] * 25;
    const float xi = xs[i], yi = ys[i];
    int32_t cnt = 0;
    for (int32_t off = 0; off < 25; ++off) {
      const int32_t nc = row[off];
      if (nc < 0) continue;
      const int32_t s = cell_start[nc], en = cell_start[nc + 1];
      if (off == 12) {
        cnt += en - s;
        if (cnt >= min_samples) break;
      } else {
        for (int32_t j = s; j < en; ++j) {
          const float ex = xs[j] - xi, ey = ys[j] - yi;
          if (ex * ex + ey * ey <= eps_sq) ++cnt;
        }
        if (cnt >= min_samples) break;
      }
    }
    if (cnt >= min_samples) is_core[i] = 1;
  }
  for (int32_t c = 0; c < num_cells; ++c) {                     /* status, :310-321 */
    const int32_t s = cell_start[c], en = cell_start[c + 1];
    int any = 0, all = (en > s);
    for (int32_t j = s; j < en; ++j) {
      if (is_core[j]) any = 1;
      else all = 0;
    }
    cell_status[c] = !any ? 0 : (all ? 2 : 1);
  }

  /* Stage 6: cell bounding boxes (dbscan_grid.hpp:332-349) and cell-level
   * union over eps-close core pairs, with the reference's prunes
   * (dbscan_grid.hpp:354-387). */
  bbox = (float*)malloc(sizeof(float) * (size_t)num_cells * 4);
  uf = (int32_t*)malloc(sizeof(int32_t) * (size_t)num_cells);
  if (!bbox || !uf) goto done;
  for (int32_t c = 0; c < num_cells; ++c) {
    const int32_t s = cell_start[c], en = cell_start[c + 1];
    float xmin = xs[s], xmax = xs[s], ymin = ys[s], ymax = ys[s];
    for (int32_t i = s + 1; i < en; ++i) {
      xmin = std_min(xmin, xs[i]);
      xmax = std_max(xmax, xs[i]);
      ymin = std_min(ymin, ys[i]);
      ymax = std_max(ymax, ys[i]);
    }
    float* box = bbox + (size_t)c * 4;
    box[0] = xmin; box[1] = xmax; box[2] = ymin; box[3] = ymax;
    uf[c] = c;
  }
  for (int32_t c = 0; c < num_cells; ++c) {
    if (cell_status[c] == 0) continue;
    const int32_t* row = ngh + (size_t)c * 25;
    const float* bb_c = bbox + (size_t)c * 4;
    const int32_t s_c = cell_start[c], e_c = cell_start[c + 1];
    for (int32_t off = 0; off < 25; ++off) {
      if (off == 12) continue;
      const int32_t nc = row[off];
      if (nc <= c) continue;
      if (cell_status[nc] == 0) continue;
      if (uf_find(uf, c) == uf_find(uf, nc)) continue;
      const float* bb_n = bbox + (size_t)nc * 4;
      if (box_box_d2(bb_c, bb_n) > eps_sq) continue;
      const int32_t s_n = cell_start[nc], e_n = cell_start[nc + 1];
      int found = 0;
      for (int32_t i = s_c; i < e_c && !found; ++i) {
        if (!is_core[i]) continue;
        const float xi = xs[i], yi = ys[i];
        if (point_box_d2(xi, yi, bb_n) > eps_sq) continue;
        for (int32_t j = s_n; j < e_n; ++j) {
          if (!is_core[j]) continue;
          const float ex = xs[j] - xi, ey = ys[j] - yi;
          if (ex * ex + ey * ey <= eps_sq) { found = 1; break; }
        }
      }
      if (found) uf_unite(uf, c, nc);
    }
  }

  /* Stage 7: dense cluster ids in cell order (dbscan_grid.hpp:392-398). */
  root_to_cluster = (int32_t*)malloc(sizeof(int32_t) * (size_t)num_cells);
  if (!root_to_cluster) goto done;
  for (int32_t c = 0; c < num_cells; ++c) root_to_cluster[c] = ORACLE_NOISE;
  int32_t cid = 0;
  for (int32_t c = 0; c < num_cells; ++c) {
    if (cell_status[c] == 0) continue;
    const int32_t r = uf_find(uf, c);
    if (root_to_cluster[r] == ORACLE_NOISE) root_to_cluster[r] = cid++;
  }

  /* Stage 8: core labels (dbscan_grid.hpp:403-412). */
  sorted_labels = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  if (!sorted_labels) goto done;
  for (int32_t i = 0; i < n; ++i) sorted_labels[i] = ORACLE_NOISE;
  for (int32_t c = 0; c < num_cells; ++c) {
    if (cell_status[c] == 0) continue;
    const int32_t lab = root_to_cluster[uf_find(uf, c)];
    for (int32_t j = cell_start[c]; j < cell_start[c + 1]; ++j)
      if (is_core[j]) sorted_labels[j] = lab;
  }

  /* Stage 9: border points take the label of the first strictly nearer core
   * within eps, scanning off = 0..24 (dbscan_grid.hpp:419-445).  Reads only
   * core labels, writes only non-core slots, so the loop is race-free. */
  const float best_init = eps_sq + 1.0f;
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t i = 0; i < n; ++i) {
    if (is_core[i]) continue;
    const int32_t* row = ngh + (size_t)point_cell[i] * 25;
    const float xi = xs[i], yi = ys[i];
    float best_d2 = best_init;
    int32_t best_lab = ORACLE_NOISE;
    for (int32_t off = 0; off < 25; ++off) {
      const int32_t nc = row[off];
      if (nc < 0 || cell_status[nc] == 0) continue;
      if (point_box_d2(xi, yi, bbox + (size_t)nc * 4) > eps_sq) continue;
      for (int32_t j = cell_start[nc]; j < cell_start[nc + 1]; ++j) {
        if (!is_core[j]) continue;
        const float ex = xs[j] - xi, ey = ys[j] - yi;
        const float d2 = ex * ex + ey * ey;
        if (d2 <= eps_sq && d2 < best_d2) {
          best_d2 = d2;
          best_lab = sorted_labels[j];
        }
      }
    }
    sorted_labels[i] = best_lab;
  }

  /* Stage 10: unpermute (dbscan_grid.hpp:450-454). */
  for (int32_t i = 0; i < n; ++i) labels[perm[i]] = sorted_labels[i];

  if (dbg) {
    dbg->num_cells = num_cells;
    dbg->n_clusters = cid;
    if (dbg->perm) memcpy(dbg->perm, perm, sizeof(int32_t) * (size_t)n);
    if (dbg->is_core) memcpy(dbg->is_core, is_core, (size_t)n);
    if (dbg->cell_start) memcpy(dbg->cell_start, cell_start, sizeof(int32_t) * ((size_t)num_cells + 1));
    if (dbg->cell_kx) memcpy(dbg->cell_kx, cell_kx, sizeof(int32_t) * (size_t)num_cells);
    if (dbg->cell_ky) memcpy(dbg->cell_ky, cell_ky, sizeof(int32_t) * (size_t)num_cells);
    if (dbg->cell_status) memcpy(dbg->cell_status, cell_status, (size_t)num_cells);
    for (int32_t c = 0; c < num_cells; ++c) {
      const int32_t r = uf_find(uf, c);
      if (dbg->cell_root) dbg->cell_root[c] = r;
      if (dbg->cell_cluster) dbg->cell_cluster[c] = cell_status[c] ? root_to_cluster[r] : ORACLE_NOISE;
    }
  }
  rc = 0;

done:
  free(keys); free(perm); free(skeys); free(xs); free(ys); free(bbox);
  free(cell_start); free(cell_kx); free(cell_ky); free(ngh); free(point_cell);
  free(uf); free(root_to_cluster); free(sorted_labels); free(is_core); free(cell_status);
  return rc;
}

/* Public entry points (called through ctypes from oracle/__init__.py). */

int oracle_dbscan2d(const float* X, int64_t n, double eps, int64_t min_samples,
                    int32_t* labels) {
  return run(X, n, eps, min_samples, labels, NULL);
}

int oracle_dbscan2d_debug(const float* X, int64_t n, double eps, int64_t min_samples,
                          int32_t* labels, int32_t* perm, uint8_t* is_core,
                          int32_t* cell_start, int32_t* cell_kx, int32_t* cell_ky,
                          uint8_t* cell_status, int32_t* cell_root,
                          int32_t* cell_cluster, int32_t* num_cells,
                          int32_t* n_clusters) {
  oracle_debug_t dbg;
  dbg.perm = perm; dbg.is_core = is_core; dbg.cell_start = cell_start;
  dbg.cell_kx = cell_kx; dbg.cell_ky = cell_ky; dbg.cell_status = cell_status;
  dbg.cell_root = cell_root; dbg.cell_cluster = cell_cluster;
  const int rc = run(X, n, eps, min_samples, labels, &dbg);
  if (num_cells) *num_cells = dbg.num_cells;
  if (n_clusters) *n_clusters = dbg.n_clusters;
  return rc;
}

int oracle_set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
  return omp_get_max_threads();
#else
  (void)threads;
  return 1;
#endif
}
