/*
 * dbscan_hip.h -- C-ABI of the MI355X (gfx950) 2-D grid DBSCAN.
 *
 * Exported by libdbscan_hip.so.  Plain pointers, sizes and a HIP stream; no
 * torch types.  The library allocates nothing: the caller passes device
 * memory for the input, the labels and one workspace buffer, and all work is
 * enqueued on `stream` without a host synchronisation (except in the profile
 * entry point, which waits for its own timing events).
 *
 * Semantics are those of the reference CPU path (csrc/cpu/dbscan_grid.hpp
 * :136-458): labels are bit-identical to it, -1 marks noise, and cluster ids
 * are numbered in the reference's cell order.
 */
#ifndef DBSCAN_HIP_H_
#define DBSCAN_HIP_H_

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes returned by every entry point. */
enum {
  DBSCAN_HIP_OK = 0,
  DBSCAN_HIP_ERR_N = 1,              /* n < 0 or n > INT32_MAX                     */
  DBSCAN_HIP_ERR_EPS = 2,            /* eps is not > 0 (NaN included)              */
  DBSCAN_HIP_ERR_MIN_SAMPLES = 3,    /* min_samples < 1                            */
  DBSCAN_HIP_ERR_NULL = 4,           /* a required pointer is NULL                 */
  DBSCAN_HIP_ERR_WORKSPACE = 5,      /* workspace smaller than the required bytes  */
  DBSCAN_HIP_ERR_ALIGN = 6,          /* X, labels or workspace misaligned          */
  DBSCAN_HIP_ERR_HIP = 7,            /* a HIP runtime call failed (see last error) */
  DBSCAN_HIP_ERR_CONFIG = 8,         /* invalid dbscan_hip_config                  */
  DBSCAN_HIP_ERR_DEVICE = 9,         /* a device-side check failed (dbscan2d_hip_check) */
  DBSCAN_HIP_ERR_TOO_LARGE = 10      /* n above DBSCAN_HIP_NOALLOC_MAX_POINTS        */
};

/*
 * Device-side error bits.  Every bounded wait inside the kernels (look-back
 * spins, the sort's pass hand-off) and the union-find step loop gives up
 * after a fixed number of steps instead of hanging, and every computed index
 * is range-checked before a write; a failure sets a bit in the workspace's
 * error word.  None is expected on valid input.  A call that ends with any
 * bit set writes INT32_MIN to every label (so no failed result passes for a
 * valid one), and the word is read back with dbscan2d_hip_error_async or
 * dbscan2d_hip_check.  The reference has no counterpart (its kernels have no
 * bounded waits; it checks cudaGetLastError after its launches,
 * csrc/cuda/dbscan_cuda.cu:738).
 */
#define DBSCAN_HIP_DEV_SORT_WAIT 1
#define DBSCAN_HIP_DEV_SCAN_WAIT 2
#define DBSCAN_HIP_DEV_UF_LOOP 4
#define DBSCAN_HIP_DEV_RANGE 8

/*
 * Launch configuration (no counterpart in the reference, whose launch shapes
 * are fixed in csrc/cuda/dbscan_cuda.cu).  Every valid configuration gives
 * the same labels; they differ in speed only.  A call (or a batch whose
 * clouds all have) at most small_max points runs as one workgroup per cloud
 * with its working set in LDS, instead of the multi-kernel pipeline; the
 * profile, timed and debug entry points always run the pipeline.  The entry points without a
 * configuration argument (or given NULL) use the built-in table for the
 * input size (dbscan2d_hip_default_config), which the project's autotuner
 * (dbscan_torch.autotune) produces and can override per process.
 */
typedef struct {
  int32_t sort_items;   /* points per thread of a radix-sort tile (512 threads): 8, 12 or 16 */
  int32_t radix_bits;   /* widest radix digit used when it saves no pass: 4..9   */
  int32_t cell_blocks;  /* workgroups (of 256) of the per-cell kernels: 64..65536 */
  int32_t pair_blocks;  /* workgroups of the cell-pair union kernels: 64..65536  */
  int32_t small_max;    /* clouds of at most this many points are clustered by
                           one workgroup in LDS (one launch); 0: never;
                           0..DBSCAN_HIP_SMALL_MAX                           */
} dbscan_hip_config;

/* Largest cloud the one-workgroup path takes (its LDS holds the whole cloud). */
#define DBSCAN_HIP_SMALL_MAX 4096

/* The built-in configuration for n points. */
int dbscan2d_hip_default_config(int64_t n, dbscan_hip_config* cfg);

/* Number of stages reported by dbscan2d_hip_profile, and their names.  The
 * names are the reference's kStageNames (csrc/cuda/dbscan_cuda.cu:513-524,
 * csrc/cpu/dbscan_grid.hpp:67-78). */
#define DBSCAN_HIP_NUM_STAGES 10

/*
 * Workspace bytes needed for n points.  Pure host arithmetic; no device call.
 * Replaces the per-call torch::empty allocations of the reference's
 * dbscan2d_cuda_impl (csrc/cuda/dbscan_cuda.cu:573-732).
 */
int dbscan2d_hip_workspace_bytes(int64_t n, size_t* bytes);

/*
 * Clusters the n points X[2*i], X[2*i+1] (float32, interleaved, device
 * memory, 8-byte aligned) and writes n int32 labels to device memory
 * `labels`.  eps and min_samples are the reference's arguments: eps is
 * validated as a double (> 0) and rounded to float, min_samples is validated
 * as >= 1 and narrowed to int32 (csrc/extension.cpp:26-31,
 * csrc/cpu/dbscan_cpu.cpp:52-53).  `workspace` (256-byte aligned) holds at
 * least dbscan2d_hip_workspace_bytes(n) bytes.
 *
 * Replaces dbscan2d_cuda (csrc/cuda/dbscan_cuda.cu:752-755), which the
 * reference's binding calls at csrc/extension.cpp:36-37.
 */
int dbscan2d_hip(const float* X, int64_t n, double eps, int64_t min_samples,
                 int32_t* labels, void* workspace, size_t workspace_bytes,
                 hipStream_t stream);

/*
 * dbscan2d_hip with an explicit launch configuration (NULL: the built-in
 * one).  No counterpart in the reference: the autotuner's entry point.
 */
int dbscan2d_hip_configured(const float* X, int64_t n, double eps, int64_t min_samples,
                            int32_t* labels, void* workspace, size_t workspace_bytes,
                            hipStream_t stream, const dbscan_hip_config* cfg);

/*
 * dbscan2d_hip_configured that also reports the call's device error word:
 * its last kernel writes the word to *error_out (device memory or pinned
 * host memory; NULL: not written), in stream order, with no extra launch.
 * Not written when n == 0 (nothing is enqueued).  No counterpart in the
 * reference; the binding uses it to check calls it does not wait for.
 */
int dbscan2d_hip_checked(const float* X, int64_t n, double eps, int64_t min_samples,
                         int32_t* labels, void* workspace, size_t workspace_bytes,
                         hipStream_t stream, const dbscan_hip_config* cfg, int32_t* error_out);

/*
 * Same as dbscan2d_hip, and additionally writes the GPU time of each of the
 * DBSCAN_HIP_NUM_STAGES stages, in milliseconds, to stage_ms (host memory).
 * Synchronises on its own events before returning.
 *
 * Replaces dbscan2d_cuda_profile (csrc/cuda/dbscan_cuda.cu:760-770), bound at
 * csrc/extension.cpp:52-63.
 */
int dbscan2d_hip_profile(const float* X, int64_t n, double eps, int64_t min_samples,
                         int32_t* labels, void* workspace, size_t workspace_bytes,
                         hipStream_t stream, float* stage_ms);

/* Name of stage i (0 <= i < DBSCAN_HIP_NUM_STAGES), or NULL. */
const char* dbscan2d_hip_stage_name(int i);

/* Number of kernel launch slots of one call (sort passes beyond the key width
 * are launched and return at once), and their names. */
#define DBSCAN_HIP_NUM_KERNELS 23
const char* dbscan2d_hip_kernel_name(int i);

/*
 * Same as dbscan2d_hip, and records events[k] on `stream` before kernel slot
 * k and events[DBSCAN_HIP_NUM_KERNELS] after the last one (the caller creates
 * the DBSCAN_HIP_NUM_KERNELS + 1 events).  Does not synchronise, so calls can
 * be timed back to back; cfg as for dbscan2d_hip_configured.  No
 * counterpart in the reference: per-kernel timing for the benchmark's
 * roofline figures and the autotuner.
 */
int dbscan2d_hip_timed(const float* X, int64_t n, double eps, int64_t min_samples,
                       int32_t* labels, void* workspace, size_t workspace_bytes,
                       hipStream_t stream, hipEvent_t* events, const dbscan_hip_config* cfg);

/*
 * Batched form: num_clouds independent clouds packed one after another in X
 * (total_points points in all, cloud i being points [offsets[i],
 * offsets[i+1])), each with its own eps[i] and min_samples[i], clustered in
 * one pipeline of launches.  labels (total_points int32, device memory) gets,
 * for every cloud, exactly the labels dbscan2d_hip would give that cloud
 * alone, numbered from 0 within the cloud.  offsets (num_clouds + 1 values,
 * offsets[0] == 0, non-decreasing, offsets[num_clouds] == total_points), eps
 * and min_samples are HOST arrays, validated per cloud as dbscan2d_hip
 * validates its arguments; on a per-cloud error *bad_cloud (if not NULL)
 * names the cloud.  total_points < INT32_MAX.  The per-cloud records and tile
 * tables are copied into the workspace with hipMemcpyAsync on `stream`.  cfg
 * as for dbscan2d_hip_configured (NULL: built-in, by total_points).
 *
 * No counterpart in the reference, whose binding clusters one cloud per call
 * (csrc/extension.cpp:33-45); this is the batched launch that replaces a
 * Python loop of dbscan2d calls over many clouds.
 */
int dbscan2d_hip_batched_workspace_bytes(int64_t total_points, int64_t num_clouds,
                                         size_t* bytes);
int dbscan2d_hip_batched(const float* X, int64_t total_points, const int64_t* offsets,
                         const double* eps, const int64_t* min_samples, int64_t num_clouds,
                         int32_t* labels, void* workspace, size_t workspace_bytes,
                         hipStream_t stream, int* bad_cloud, const dbscan_hip_config* cfg);

/* dbscan2d_hip_batched that also reports the batch's device error word, as
 * dbscan2d_hip_checked does.  No counterpart in the reference. */
int dbscan2d_hip_batched_checked(const float* X, int64_t total_points, const int64_t* offsets,
                                 const double* eps, const int64_t* min_samples,
                                 int64_t num_clouds, int32_t* labels, void* workspace,
                                 size_t workspace_bytes, hipStream_t stream, int* bad_cloud,
                                 const dbscan_hip_config* cfg, int32_t* error_out);

/*
 * dbscan2d_hip_batched in two steps, so that the launches can be captured in
 * a HIP graph once per batch shape and replayed (no reference counterpart).
 *
 * dbscan2d_hip_batched_stage validates the batch as dbscan2d_hip_batched
 * does, copies its per-cloud records and tile tables into `workspace` with
 * hipMemcpyAsync on `stream` (the host arrays may be reused when it
 * returns), and fills *plan with what the launches need.
 * dbscan2d_hip_batched_enqueue then enqueues kernels only -- no copy,
 * allocation, memset or host synchronisation -- on `stream`: the clustering
 * of the points at X into labels (both device memory, as for
 * dbscan2d_hip_batched), with the staged tables; error_out as for
 * dbscan2d_hip_batched_checked.  It may be captured with
 * hipStreamBeginCapture and the graph replayed any number of times: each
 * run starts from the staged tables, so a replay after X is refilled with
 * other clouds of the same sizes, eps and min_samples gives their labels.
 * Restage (same workspace) to change eps or min_samples but not the sizes;
 * a batch of other sizes needs its own stage, plan and graph.  Runs that
 * share a workspace must not overlap.  Labels equal dbscan2d_hip_batched's.
 */
typedef struct {
  int64_t total_points, num_clouds;
  int64_t sort_tiles, scan_tiles;  /* launch sizes */
  int32_t all_small;               /* 1: the one-workgroup persistent grid */
  int32_t reserved_;
  dbscan_hip_config cfg;
} dbscan_hip_batch_plan;

int dbscan2d_hip_batched_stage(int64_t total_points, const int64_t* offsets, const double* eps,
                               const int64_t* min_samples, int64_t num_clouds, void* workspace,
                               size_t workspace_bytes, hipStream_t stream, int* bad_cloud,
                               const dbscan_hip_config* cfg, dbscan_hip_batch_plan* plan);
int dbscan2d_hip_batched_enqueue(const float* X, int32_t* labels, void* workspace,
                                 size_t workspace_bytes, hipStream_t stream,
                                 const dbscan_hip_batch_plan* plan, int32_t* error_out);

/*
 * dbscan2d_hip_batched with the event slots of dbscan2d_hip_timed (slot 0
 * also covers the per-cloud key-map kernel).  No counterpart in the
 * reference: per-kernel timing of batches.
 */
int dbscan2d_hip_batched_timed(const float* X, int64_t total_points, const int64_t* offsets,
                               const double* eps, const int64_t* min_samples,
                               int64_t num_clouds, int32_t* labels, void* workspace,
                               size_t workspace_bytes, hipStream_t stream,
                               hipEvent_t* events, const dbscan_hip_config* cfg);

/*
 * Byte offsets, inside the workspace of a dbscan2d_hip call on n points, of
 * the intermediates that call leaves behind (all in the library's sorted
 * point order and internal cell order).  For stage-wise comparison with the
 * oracle; replaces the intermediates of dbscan2d_cuda_debug
 * (csrc/cuda/dbscan_cuda.cu:777-800).  Fields are filled for `n` points; the
 * counts that are only known on the device (cells, sort passes) are read from
 * the `info` block.
 */
typedef struct {
  size_t info;        /* int32[8]: kxmin, kxmax, kymin, kymax, num_cells, npasses,
                         error bits (0 when every bounded wait and index check
                         passed), column-index entries of a batch (0) */
  size_t xs_a, ys_a, idx_a;   /* sorted x, y, original index if npasses is odd  */
  size_t xs_b, ys_b, idx_b;   /* ... if npasses is even                         */
  size_t point_cell;  /* int32[n]  cell of each sorted point                      */
  size_t is_core;     /* uint8[n]  core flag of each sorted point                 */
  size_t cell_start;  /* int32[num_cells+1]                                       */
  size_t cell_key;    /* uint64[num_cells]  (x' << 32 | y'), internal coordinates */
  size_t cell_flags;  /* uint32[num_cells]  bit0 has core, bit1 has non-core      */
  size_t cell_ref;    /* int32[num_cells]   reference-order index of each cell    */
  size_t uf;          /* int32[num_cells]   union-find root, by reference index   */
  size_t cell_cluster;/* int32[num_cells]   cluster id, by internal index, or -1  */
} dbscan_hip_layout;

int dbscan2d_hip_layout(int64_t n, dbscan_hip_layout* layout);

/*
 * Allocation-free variant of dbscan2d_hip, with no counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * takes no workspace: the pipeline runs in a file-scope device array of
 * DBSCAN_HIP_NOALLOC_SCRATCH_BYTES bytes that each GPU the library runs on
 * holds until the process ends; its device address is looked up once per
 * device on the first call there.  The call allocates nothing and enqueues
 * only kernels on `stream`.  n may be at most DBSCAN_HIP_NOALLOC_MAX_POINTS
 * (DBSCAN_HIP_ERR_TOO_LARGE above it).  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (on different streams, or from
 * different threads without ordering) race and must not be made; calls
 * ordered on one stream are safe.  Each call clears what it needs, so the
 * previous call's contents do not matter.  A failed device-side check leaves
 * every label INT32_MIN.  Labels equal dbscan2d_hip's.
 */
#define DBSCAN_HIP_NOALLOC_SCRATCH_BYTES (160u << 20)
#define DBSCAN_HIP_NOALLOC_MAX_POINTS (1 << 20)
int dbscan2d_hip_noalloc(const float* X, int64_t n, double eps, int64_t min_samples,
                         int32_t* labels, hipStream_t stream);

/* Human-readable text for a status code. */
const char* dbscan2d_hip_status_string(int status);

/* The hipError_t behind the most recent DBSCAN_HIP_ERR_HIP on this thread. */
int dbscan2d_hip_last_hip_error(void);

/*
 * Enqueues on `stream` a copy of the error word that the last call using
 * `workspace` (single or batched) leaves, into *dst: device memory or pinned
 * (hipHostMalloc) host memory, valid once the stream has passed this point.
 * One one-thread kernel; no synchronisation.  Pairs with the C10_CUDA_CHECK /
 * cudaGetLastError checks of the reference binding (csrc/extension.cpp:36-42,
 * csrc/cuda/dbscan_cuda.cu:738), which see launch errors only.
 */
int dbscan2d_hip_error_async(const void* workspace, int32_t* dst, hipStream_t stream);

/*
 * Waits for `stream`, then reads the error word of the last call that used
 * `workspace`: DBSCAN_HIP_OK if it is 0, else DBSCAN_HIP_ERR_DEVICE with the
 * word in *bits (if not NULL).
 */
int dbscan2d_hip_check(const void* workspace, hipStream_t stream, int32_t* bits);

/* Text for one device-side error bit (DBSCAN_HIP_DEV_*), or NULL. */
const char* dbscan2d_hip_device_error_name(int32_t bit);

/*
 * Test hook (no counterpart in the reference): the next call on this host
 * thread that runs the pipeline starts with `bits` in its error word, so the
 * reporting path (poisoned labels, error word, exceptions) can be tested
 * without making any kernel fail.  Consumed by that call.
 */
void dbscan2d_hip_inject_error_for_testing(int32_t bits);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* DBSCAN_HIP_H_ */
