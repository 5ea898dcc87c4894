// Python-facing entry points of dbscan_torch on PyTorch-ROCm.
//
// Mirrors the reference binding (csrc/extension.cpp:26-81): the same module
// name (_C), function names, argument names, validation order and error
// messages.  The compute arm is the gfx950 HIP pipeline behind the C-ABI in
// include/dbscan_hip.h; this file only validates, stages memory through torch's
// caching allocator, and passes raw device pointers and the current HIP stream.
//
// Host (CPU) tensors are clustered on the current HIP device and the labels
// are copied back, so the result stays on X's device as the reference promises
// (csrc/extension.cpp:33-45).  There is no CPU compute path: without a HIP
// device a host tensor raises.

#include <torch/extension.h>

#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPFunctions.h>
#include <c10/hip/HIPStream.h>

#include <algorithm>
#include <deque>
#include <map>
#include <mutex>
#include <optional>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "dbscan_hip.h"

namespace {

// csrc/extension.cpp:26-31 -- same checks, same order, same messages.
void check_input(const torch::Tensor& X, double eps, int64_t min_samples) {
  TORCH_CHECK(X.dim() == 2 && X.size(1) == 2, "X must have shape (N, 2)");
  TORCH_CHECK(X.scalar_type() == torch::kFloat32, "X must be float32");
  TORCH_CHECK(eps > 0.0, "eps must be > 0");
  TORCH_CHECK(min_samples >= 1, "min_samples must be >= 1");
}

void check_status(int rc) {
  if (rc == DBSCAN_HIP_ERR_HIP) {
    TORCH_CHECK(false, "dbscan2d: HIP error ", dbscan2d_hip_last_hip_error(), " (",
                hipGetErrorString(static_cast<hipError_t>(dbscan2d_hip_last_hip_error())), ")");
  }
  TORCH_CHECK(rc == DBSCAN_HIP_OK, "dbscan2d: ", dbscan2d_hip_status_string(rc));
}

// Text of a non-zero device error word: each set bit named.
std::string device_error_text(int32_t bits) {
  std::string why;
  for (int b = 0; b < 31; ++b) {
    const int32_t bit = static_cast<int32_t>(1u << b);
    if (!(bits & bit)) continue;
    const char* name = dbscan2d_hip_device_error_name(bit);
    if (!why.empty()) why += "; ";
    why += name ? name : ("unknown bit " + std::to_string(b));
  }
  char hex[16];
  std::snprintf(hex, sizeof(hex), "0x%x", static_cast<unsigned>(bits));
  return std::string("device-side check failed (error bits ") + hex + ": " + why +
         "); every label of that call is INT32_MIN";
}

void raise_for_device_error(int32_t bits, const char* which = "dbscan2d: ") {
  TORCH_CHECK(bits == 0, which, device_error_text(bits));
}

// Error words of calls that returned device labels without a host
// synchronisation: each such call has its last kernel write its word into a
// pinned slot (dbscan2d_hip_checked) and records an event behind it; the
// words of finished calls are examined, oldest first, at the start of every
// later call on the device and by check_device_errors(), which raise for a
// non-zero word.  (A call whose labels come back to the host checks its own
// word before returning.)
struct ErrorWatch {
  static constexpr int kSlots = 64;
  int32_t* words = nullptr;  // pinned host memory, one word per slot
  hipEvent_t ev[kSlots] = {};
  std::deque<int> pending;   // slots in enqueue order
  int next = 0;
};
std::mutex g_watch_mu;
std::map<int, ErrorWatch> g_watch;

ErrorWatch& watch_for(int device) {
  ErrorWatch& w = g_watch[device];
  if (!w.words) {
    void* p = nullptr;
    C10_HIP_CHECK(hipHostMalloc(&p, sizeof(int32_t) * ErrorWatch::kSlots, hipHostMallocDefault));
    w.words = static_cast<int32_t*>(p);
    for (auto& e : w.ev) C10_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  return w;
}

// True while `stream` is being captured into a HIP graph: event queries and
// waits are not allowed then, so captured calls are neither polled nor
// watched (a replayed call that fails a check still leaves its labels
// INT32_MIN).
bool capturing(hipStream_t stream) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(stream, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// Examines the pending words of `device` (waiting for them if `wait`);
// raises for the first non-zero one.
void poll_device_errors(int device, bool wait) {
  int32_t bits = 0;
  {
    std::lock_guard<std::mutex> lock(g_watch_mu);
    auto it = g_watch.find(device);
    if (it == g_watch.end()) return;
    ErrorWatch& w = it->second;
    while (!w.pending.empty() && bits == 0) {
      const int slot = w.pending.front();
      if (wait) {
        C10_HIP_CHECK(hipEventSynchronize(w.ev[slot]));
      } else {
        const hipError_t q = hipEventQuery(w.ev[slot]);
        if (q == hipErrorNotReady) break;
        C10_HIP_CHECK(q);
      }
      w.pending.pop_front();
      bits = w.words[slot];
    }
  }
  raise_for_device_error(bits, "dbscan2d: an earlier call's ");
}

// A pinned slot for the error word of a call about to be enqueued on
// `device` (the oldest slot is waited for and examined if all are pending).
struct WatchSlot {
  int device = -1, slot = -1;
  int32_t* word = nullptr;
};

WatchSlot watch_begin(int device) {
  int32_t bits = 0;
  WatchSlot ws;
  {
    std::lock_guard<std::mutex> lock(g_watch_mu);
    ErrorWatch& w = watch_for(device);
    if (static_cast<int>(w.pending.size()) == ErrorWatch::kSlots) {
      const int old = w.pending.front();
      C10_HIP_CHECK(hipEventSynchronize(w.ev[old]));
      w.pending.pop_front();
      bits = w.words[old];
    }
    ws.device = device;
    ws.slot = w.next;
    w.next = (w.next + 1) % ErrorWatch::kSlots;
    ws.word = w.words + ws.slot;
  }
  raise_for_device_error(bits, "dbscan2d: an earlier call's ");
  return ws;
}

// Marks the slot's word as due once `stream` reaches this point.
void watch_end(const WatchSlot& ws, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(g_watch_mu);
  ErrorWatch& w = g_watch[ws.device];
  C10_HIP_CHECK(hipEventRecord(w.ev[ws.slot], stream));
  w.pending.push_back(ws.slot);
}

// Waits for `stream` and raises if the call that last used `workspace` failed
// a device-side check.
void check_call(const void* workspace, hipStream_t stream) {
  int32_t bits = 0;
  const int rc = dbscan2d_hip_check(workspace, stream, &bits);
  if (rc != DBSCAN_HIP_ERR_DEVICE) check_status(rc);
  raise_for_device_error(bits);
}

// Launch configurations by input size, installed by the autotuner
// (dbscan_torch.autotune); empty: the library's built-in table.
std::vector<std::pair<int64_t, dbscan_hip_config>> g_tuning;

const dbscan_hip_config* config_for(int64_t n) {
  for (const auto& e : g_tuning)
    if (n <= e.first) return &e.second;
  return g_tuning.empty() ? nullptr : &g_tuning.back().second;
}

// (sort_items, radix_bits, cell_blocks, pair_blocks[, small_max]); without
// small_max, the library's default for that size.
dbscan_hip_config to_config(const std::vector<int64_t>& v, int64_t n = 0) {
  TORCH_CHECK(v.size() == 4 || v.size() == 5,
              "config must be (sort_items, radix_bits, cell_blocks, pair_blocks[, small_max])");
  dbscan_hip_config d;
  check_status(dbscan2d_hip_default_config(n, &d));
  return dbscan_hip_config{static_cast<int32_t>(v[0]), static_cast<int32_t>(v[1]),
                           static_cast<int32_t>(v[2]), static_cast<int32_t>(v[3]),
                           v.size() == 5 ? static_cast<int32_t>(v[4]) : d.small_max};
}

// The configuration a call on n points uses, with the one-workgroup path
// switched off (the debug variants read the pipeline's intermediates).
dbscan_hip_config pipeline_config(int64_t n) {
  dbscan_hip_config c;
  const dbscan_hip_config* t = config_for(n);
  if (t) c = *t;
  else check_status(dbscan2d_hip_default_config(n, &c));
  c.small_max = 0;
  return c;
}

struct DeviceRun {
  torch::Tensor Xd;         // contiguous input on the HIP device
  torch::Tensor labels;     // int32 labels on the HIP device
  torch::Tensor workspace;  // uint8 workspace (kept alive by the debug binding)
};

// How a call's device-side error word is checked: before returning (the
// result goes back to the host, or the call waits anyway), by the watch
// (device results), or not at all (the debug variant returns the word).
enum class Check { kNow, kWatch, kNone };

Check check_for(const torch::Tensor& X, bool waits) {
  return (!X.is_cuda() || waits) ? Check::kNow : Check::kWatch;
}

// Runs the pipeline; stage_ms may be null.  Returns device tensors.
DeviceRun run_device(const torch::Tensor& X, double eps, int64_t min_samples, float* stage_ms,
                     const dbscan_hip_config* cfg = nullptr, Check check = Check::kWatch) {
  c10::Device device = X.device();
  if (!X.is_cuda()) {
    TORCH_CHECK(c10::hip::device_count() > 0,
                "dbscan2d: no HIP device is available. dbscan_torch clusters on an AMD GPU "
                "(gfx950); pass a tensor on a ROCm device or run on a machine with one.");
    device = c10::Device(c10::kCUDA, c10::hip::current_device());
  }
  c10::DeviceGuard guard(device);
  hipStream_t stream = c10::hip::getCurrentHIPStream(device.index()).stream();
  const bool in_graph = capturing(stream);
  if (!in_graph) poll_device_errors(device.index(), false);
  DeviceRun r;
  r.Xd = X.to(device).contiguous();
  const int64_t n = r.Xd.size(0);
  TORCH_CHECK(n <= INT32_MAX, "dbscan2d: at most 2^31 - 1 points are supported");
  size_t ws_bytes = 0;
  check_status(dbscan2d_hip_workspace_bytes(n, &ws_bytes));
  auto opts = torch::TensorOptions().device(device);
  r.workspace = torch::empty({static_cast<int64_t>(ws_bytes)}, opts.dtype(torch::kUInt8));
  r.labels = torch::empty({n}, opts.dtype(torch::kInt32));
  const bool watch = check == Check::kWatch && !in_graph && !stage_ms;
  WatchSlot slot;
  if (watch) slot = watch_begin(device.index());
  const int rc =
      stage_ms ? dbscan2d_hip_profile(r.Xd.data_ptr<float>(), n, eps, min_samples,
                                      r.labels.data_ptr<int32_t>(), r.workspace.data_ptr(),
                                      ws_bytes, stream, stage_ms)
               : dbscan2d_hip_checked(r.Xd.data_ptr<float>(), n, eps, min_samples,
                                      r.labels.data_ptr<int32_t>(), r.workspace.data_ptr(),
                                      ws_bytes, stream, cfg ? cfg : config_for(n), slot.word);
  check_status(rc);
  if (watch) watch_end(slot, stream);
  else if (check == Check::kNow && !in_graph) check_call(r.workspace.data_ptr(), stream);
  return r;
}

torch::Tensor finish(const torch::Tensor& X, const torch::Tensor& labels) {
  return X.is_cuda() ? labels : labels.to(X.device());
}

}  // namespace

// csrc/extension.cpp:33-45
torch::Tensor dbscan2d(torch::Tensor X, double eps, int64_t min_samples) {
  check_input(X, eps, min_samples);
  if (X.size(0) == 0) {
    return torch::empty({0}, X.options().dtype(torch::kInt32));
  }
  return finish(X, run_device(X, eps, min_samples, nullptr, nullptr, check_for(X, false)).labels);
}

// dbscan2d with an explicit launch configuration (no reference counterpart;
// for the autotuner and its parity tests).
torch::Tensor dbscan2d_configured(torch::Tensor X, double eps, int64_t min_samples,
                                  std::vector<int64_t> config) {
  check_input(X, eps, min_samples);
  if (X.size(0) == 0) return torch::empty({0}, X.options().dtype(torch::kInt32));
  const dbscan_hip_config cfg = to_config(config, X.size(0));
  return finish(X, run_device(X, eps, min_samples, nullptr, &cfg, check_for(X, false)).labels);
}

void set_tuning_table(std::vector<std::vector<int64_t>> rows) {
  std::vector<std::pair<int64_t, dbscan_hip_config>> t;
  for (const auto& r : rows) {
    TORCH_CHECK(r.size() == 5 || r.size() == 6,
                "tuning rows are (max_n, sort_items, radix_bits, cell_blocks, pair_blocks"
                "[, small_max])");
    t.emplace_back(r[0], to_config(std::vector<int64_t>(r.begin() + 1, r.end()), r[0]));
  }
  std::sort(t.begin(), t.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  g_tuning = std::move(t);
}

std::vector<std::vector<int64_t>> get_tuning_table() {
  std::vector<std::vector<int64_t>> out;
  for (const auto& e : g_tuning)
    out.push_back({e.first, e.second.sort_items, e.second.radix_bits, e.second.cell_blocks,
                   e.second.pair_blocks, e.second.small_max});
  return out;
}

std::vector<int64_t> default_config(int64_t n) {
  dbscan_hip_config c;
  check_status(dbscan2d_hip_default_config(n, &c));
  return {c.sort_items, c.radix_bits, c.cell_blocks, c.pair_blocks, c.small_max};
}

// csrc/extension.cpp:52-63: labels plus GPU milliseconds per reference stage.
std::pair<torch::Tensor, std::map<std::string, float>> dbscan2d_profile(torch::Tensor X,
                                                                        double eps,
                                                                        int64_t min_samples) {
  check_input(X, eps, min_samples);
  float ms[DBSCAN_HIP_NUM_STAGES] = {};
  torch::Tensor labels;
  if (X.size(0) == 0) {
    labels = torch::empty({0}, X.options().dtype(torch::kInt32));
  } else {
    labels = finish(X, run_device(X, eps, min_samples, ms, nullptr, Check::kNow).labels);
  }
  std::map<std::string, float> out;
  for (int i = 0; i < DBSCAN_HIP_NUM_STAGES; ++i) out[dbscan2d_hip_stage_name(i)] = ms[i];
  return {labels, out};
}

// Debug variant (the reference's dbscan2d_cuda_debug, dbscan_cuda.cu:777-800):
// returns the labels and the intermediates the pipeline leaves in its
// workspace, as device tensors in the library's internal point and cell order:
//   0 labels (n,) int32             1 info (8,) int32: kxmin, kxmax, kymin, kymax,
//                                     num_cells, npasses, error, 0
//   2 perm (n,) int32 (sorted -> original index)
//   3 xs (n,) float32   4 ys (n,) float32   (sorted order)
//   5 point_cell (n,) int32          6 is_core (n,) uint8
//   7 cell_start (C+1,) int32        8 cell_key (C,) int64 (x' << 32 | y')
//   9 cell_flags (C,) int32 (bit0 has core, bit1 has non-core)
//  10 cell_ref (C,) int32 (reference cell order index)
//  11 uf (C,) int32 (root, by reference index)
//  12 cell_cluster (C,) int32
std::vector<torch::Tensor> dbscan2d_hip_debug(torch::Tensor X, double eps, int64_t min_samples) {
  check_input(X, eps, min_samples);
  TORCH_CHECK(X.size(0) > 0, "debug variant requires n > 0");
  const dbscan_hip_config cfg = pipeline_config(X.size(0));
  DeviceRun r = run_device(X, eps, min_samples, nullptr, &cfg, Check::kNone);
  const int64_t n = r.Xd.size(0);
  dbscan_hip_layout L;
  check_status(dbscan2d_hip_layout(n, &L));
  auto ws = r.workspace;
  auto slice = [&](size_t off, int64_t count, int64_t elem, torch::ScalarType t) {
    return ws.narrow(0, static_cast<int64_t>(off), count * elem).view(t);
  };
  auto info = slice(L.info, 8, 4, torch::kInt32).clone();
  const auto info_h = info.cpu();
  const int64_t C = info_h[4].item<int64_t>();
  const int64_t npasses = info_h[5].item<int64_t>();
  const bool in_a = (npasses & 1) != 0;
  std::vector<torch::Tensor> out;
  out.push_back(r.labels);
  out.push_back(info);
  out.push_back(slice(in_a ? L.idx_a : L.idx_b, n, 4, torch::kInt32).clone());
  out.push_back(slice(in_a ? L.xs_a : L.xs_b, n, 4, torch::kFloat32).clone());
  out.push_back(slice(in_a ? L.ys_a : L.ys_b, n, 4, torch::kFloat32).clone());
  out.push_back(slice(L.point_cell, n, 4, torch::kInt32).clone());
  out.push_back(slice(L.is_core, n, 1, torch::kUInt8).clone());
  out.push_back(slice(L.cell_start, C + 1, 4, torch::kInt32).clone());
  out.push_back(slice(L.cell_key, C, 8, torch::kInt64).clone());
  out.push_back(slice(L.cell_flags, C, 4, torch::kInt32).clone());
  out.push_back(slice(L.cell_ref, C, 4, torch::kInt32).clone());
  out.push_back(slice(L.uf, C, 4, torch::kInt32).clone());
  out.push_back(slice(L.cell_cluster, C, 4, torch::kInt32).clone());
  return out;
}

// The reference's debug entry point (csrc/cuda/dbscan_cuda.cu:772-800,
// bound at csrc/extension.cpp:76-79): labels and ten stage intermediates, in
// the reference's order and in the reference CPU path's point and cell order
// (cells by the signed int64 key, points of a cell in input order; the
// library's internal cell order differs inside a column, and each cell's
// reference index maps between the two):
//   0 labels (n,) int32                    1 keys_sorted (n,) int64, the signed
//                                            key u32(kx) << 32 | u32(ky) per point
//   2 perm (n,) int32 (sorted -> input)    3 xs (n,) float32   4 ys (n,) float32
//   5 is_core (n,) uint8                   6 cell_start (C+1,) int32
//   7 cluster_id_per_cell (C,) int32       8 uf (C,) int32, each cell's root
//   9 sorted_labels_pre_border (n,) int32, core labels and -1 elsewhere
//  10 sorted_labels_final (n,) int32, after border assignment
// Device tensors; no error is raised for a failed device-side check (the
// labels are then INT32_MIN).
std::vector<torch::Tensor> dbscan2d_cuda_debug(torch::Tensor X, double eps, int64_t min_samples) {
  check_input(X, eps, min_samples);
  TORCH_CHECK(X.size(0) > 0, "debug variant requires n > 0");
  TORCH_CHECK(static_cast<int32_t>(static_cast<uint32_t>(static_cast<uint64_t>(min_samples))) >= 1,
              "debug variant: min_samples narrows below 1 (every label is -1, no stage runs)");
  const dbscan_hip_config cfg = pipeline_config(X.size(0));
  DeviceRun r = run_device(X, eps, min_samples, nullptr, &cfg, Check::kNone);
  const int64_t n = r.Xd.size(0);
  dbscan_hip_layout L;
  check_status(dbscan2d_hip_layout(n, &L));
  auto ws = r.workspace;
  auto slice = [&](size_t off, int64_t count, int64_t elem, torch::ScalarType t) {
    return ws.narrow(0, static_cast<int64_t>(off), count * elem).view(t);
  };
  const auto info = slice(L.info, 8, 4, torch::kInt32).cpu();
  const int64_t kxmin = info[0].item<int64_t>(), kymin = info[2].item<int64_t>();
  const int64_t C = info[4].item<int64_t>(), npasses = info[5].item<int64_t>();
  const bool in_a = (npasses & 1) != 0;
  const auto i64 = torch::kInt64;
  auto perm = slice(in_a ? L.idx_a : L.idx_b, n, 4, torch::kInt32);
  auto xs = slice(in_a ? L.xs_a : L.xs_b, n, 4, torch::kFloat32);
  auto ys = slice(in_a ? L.ys_a : L.ys_b, n, 4, torch::kFloat32);
  auto lab_sorted = slice(in_a ? L.idx_b : L.idx_a, n, 4, torch::kInt32);  // border labels
  auto point_cell = slice(L.point_cell, n, 4, torch::kInt32).to(i64);
  auto is_core = slice(L.is_core, n, 1, torch::kUInt8);
  auto cell_start = slice(L.cell_start, C + 1, 4, torch::kInt32).to(i64);
  auto cell_key = slice(L.cell_key, C, 8, torch::kInt64);
  auto cell_ref = slice(L.cell_ref, C, 4, torch::kInt32).to(i64);
  auto uf = slice(L.uf, C, 4, torch::kInt32).clone();
  auto cell_cluster = slice(L.cell_cluster, C, 4, torch::kInt32);
  // reference cell order: sizes by reference index, their running sum
  auto sizes = cell_start.narrow(0, 1, C) - cell_start.narrow(0, 0, C);
  auto sizes_ref = torch::empty_like(sizes).index_put_({cell_ref}, sizes);
  auto ref_start = torch::cat({torch::zeros({1}, sizes.options()), sizes_ref.cumsum(0)});
  auto pos = torch::arange(n, point_cell.options());
  auto new_pos = ref_start.index({cell_ref.index({point_cell})}) + pos -
                 cell_start.index({point_cell});
  auto remap = [&](const torch::Tensor& a) {
    return torch::empty_like(a).index_put_({new_pos}, a);
  };
  // the signed key of each cell: kx * 2^32 + u32(ky)
  const int64_t m32 = 0xFFFFFFFFll;
  auto low32 = [&](const torch::Tensor& t) { return torch::bitwise_and(t, m32); };
  auto kx_u = low32(low32(torch::bitwise_right_shift(cell_key, 32)) + (kxmin & m32));
  auto ky_u = low32(low32(cell_key) + (kymin & m32));
  auto kx_s = kx_u - (kx_u >= (int64_t{1} << 31)).to(i64) * (int64_t{1} << 32);
  auto key = kx_s * (int64_t{1} << 32) + ky_u;
  auto core = is_core.to(torch::kBool);
  auto core_lab = cell_cluster.index({point_cell});
  auto minus1 = torch::full_like(core_lab, -1);
  auto pre = torch::where(core, core_lab, minus1);
  auto fin = torch::where(core, core_lab, lab_sorted);
  auto cc_ref = torch::empty_like(cell_cluster).index_put_({cell_ref}, cell_cluster);
  return {r.labels,
          remap(key.index({point_cell})),
          remap(perm),
          remap(xs),
          remap(ys),
          remap(is_core),
          ref_start.to(torch::kInt32),
          cc_ref,
          uf,
          remap(pre),
          remap(fin)};
}

void check_device_errors(std::optional<int64_t> device) {
  if (device) {
    poll_device_errors(static_cast<int>(*device), true);
    return;
  }
  std::vector<int> devs;
  {
    std::lock_guard<std::mutex> lock(g_watch_mu);
    for (const auto& kv : g_watch) devs.push_back(kv.first);
  }
  for (int d : devs) poll_device_errors(d, true);
}

namespace {

// Checks a batch and returns its host arrays (no reference counterpart; the
// per-cloud checks and messages are the single call's, naming the cloud).
struct Batch {
  std::vector<int64_t> offsets, ms;
  std::vector<double> eps;
};

Batch check_batch(const torch::Tensor& X, const std::vector<int64_t>& offsets,
                  const std::vector<double>& eps, const std::vector<int64_t>& min_samples) {
  TORCH_CHECK(X.dim() == 2 && X.size(1) == 2, "X must have shape (N, 2)");
  TORCH_CHECK(X.scalar_type() == torch::kFloat32, "X must be float32");
  const size_t B = eps.size();
  TORCH_CHECK(min_samples.size() == B, "dbscan2d_batched: eps and min_samples differ in length");
  TORCH_CHECK(offsets.size() == B + 1, "dbscan2d_batched: offsets must hold one more value than eps");
  TORCH_CHECK(offsets.front() == 0 && offsets.back() == X.size(0),
              "dbscan2d_batched: offsets must run from 0 to N");
  for (size_t i = 0; i < B; ++i) {
    TORCH_CHECK(offsets[i + 1] >= offsets[i], "dbscan2d_batched: offsets must be non-decreasing");
    TORCH_CHECK(eps[i] > 0.0, "eps must be > 0 (cloud ", i, ")");
    TORCH_CHECK(min_samples[i] >= 1, "min_samples must be >= 1 (cloud ", i, ")");
  }
  TORCH_CHECK(X.size(0) < INT32_MAX, "dbscan2d_batched: fewer than 2^31 - 1 points in all");
  return Batch{offsets, min_samples, eps};
}

}  // namespace

// Batched clustering (no reference counterpart): every cloud i = rows
// [offsets[i], offsets[i+1]) of X is clustered with its own eps[i] and
// min_samples[i] in one pipeline; the labels of each cloud equal those of a
// dbscan2d call on that cloud alone.
torch::Tensor dbscan2d_batched(torch::Tensor X, std::vector<int64_t> offsets,
                               std::vector<double> eps, std::vector<int64_t> min_samples) {
  const Batch b = check_batch(X, offsets, eps, min_samples);
  const int64_t n = X.size(0), B = static_cast<int64_t>(eps.size());
  if (n == 0) return torch::empty({0}, X.options().dtype(torch::kInt32));
  c10::Device device = X.device();
  if (!X.is_cuda()) {
    TORCH_CHECK(c10::hip::device_count() > 0,
                "dbscan2d: no HIP device is available. dbscan_torch clusters on an AMD GPU "
                "(gfx950); pass a tensor on a ROCm device or run on a machine with one.");
    device = c10::Device(c10::kCUDA, c10::hip::current_device());
  }
  c10::DeviceGuard guard(device);
  hipStream_t stream = c10::hip::getCurrentHIPStream(device.index()).stream();
  const bool in_graph = capturing(stream);
  if (!in_graph) poll_device_errors(device.index(), false);
  auto Xd = X.to(device).contiguous();
  size_t ws_bytes = 0;
  check_status(dbscan2d_hip_batched_workspace_bytes(n, B, &ws_bytes));
  auto opts = torch::TensorOptions().device(device);
  auto workspace = torch::empty({static_cast<int64_t>(ws_bytes)}, opts.dtype(torch::kUInt8));
  auto labels = torch::empty({n}, opts.dtype(torch::kInt32));
  int bad = -1;
  const bool watch = X.is_cuda() && !in_graph;
  WatchSlot slot;
  if (watch) slot = watch_begin(device.index());
  check_status(dbscan2d_hip_batched_checked(Xd.data_ptr<float>(), n, b.offsets.data(),
                                            b.eps.data(), b.ms.data(), B,
                                            labels.data_ptr<int32_t>(), workspace.data_ptr(),
                                            ws_bytes, stream, &bad, config_for(n), slot.word));
  if (watch)
    watch_end(slot, stream);
  else if (!in_graph)
    check_call(workspace.data_ptr(), stream);
  return finish(X, labels);
}

// Serving helper (no reference counterpart): runs a batch into buffers the
// caller owns -- X (N, 2) float32, labels (>= N,) int32 and a uint8 workspace,
// all on one ROCm device, typically views into a pooled arena -- on the
// current stream, without allocating.  Returns nothing; the work is enqueued.
void dbscan2d_batched_out(torch::Tensor X, std::vector<int64_t> offsets, std::vector<double> eps,
                          std::vector<int64_t> min_samples, torch::Tensor labels,
                          torch::Tensor workspace, std::optional<torch::Tensor> error_word) {
  const Batch b = check_batch(X, offsets, eps, min_samples);
  const int64_t n = X.size(0), B = static_cast<int64_t>(eps.size());
  TORCH_CHECK(X.is_cuda() && labels.is_cuda() && workspace.is_cuda() &&
                  labels.device() == X.device() && workspace.device() == X.device(),
              "dbscan2d_batched_out: X, labels and workspace must be on one ROCm device");
  TORCH_CHECK(X.is_contiguous() && labels.is_contiguous() && workspace.is_contiguous(),
              "dbscan2d_batched_out: buffers must be contiguous");
  TORCH_CHECK(labels.scalar_type() == torch::kInt32 && labels.numel() >= n,
              "dbscan2d_batched_out: labels must be int32 with at least N elements");
  TORCH_CHECK(workspace.scalar_type() == torch::kUInt8,
              "dbscan2d_batched_out: workspace must be uint8");
  int32_t* err = nullptr;
  if (error_word) {
    const torch::Tensor& e = *error_word;
    TORCH_CHECK(e.scalar_type() == torch::kInt32 && e.numel() >= 1 && e.is_contiguous() &&
                    (e.is_cuda() || e.is_pinned()),
                "dbscan2d_batched_out: error_word must be an int32 tensor in device or pinned "
                "host memory");
    err = e.data_ptr<int32_t>();
    if (n == 0) *err = 0;  // nothing is enqueued (a device word is left as it is)
  }
  if (n == 0) return;
  c10::DeviceGuard guard(X.device());
  hipStream_t stream = c10::hip::getCurrentHIPStream(X.device().index()).stream();
  int bad = -1;
  check_status(dbscan2d_hip_batched_checked(
      X.data_ptr<float>(), n, b.offsets.data(), b.eps.data(), b.ms.data(), B,
      labels.data_ptr<int32_t>(), workspace.data_ptr(), static_cast<size_t>(workspace.numel()),
      stream, &bad, config_for(n), err));
}

// dbscan2d_batched in two steps, for HIP graph capture (no reference
// counterpart).  dbscan2d_batched_stage validates the batch (X gives the
// shape and device), copies its tables into `workspace` on the current
// stream and returns the launch plan as ten integers; dbscan2d_batched_enqueue
// enqueues the kernels only, on the current stream (capturable), into
// `labels`, writing the device error word to `error_word` if given.
namespace {
constexpr size_t kPlanInts = 10;
std::vector<int64_t> plan_to_ints(const dbscan_hip_batch_plan& p) {
  return {p.total_points, p.num_clouds,   p.sort_tiles,     p.scan_tiles,     p.all_small,
          p.cfg.sort_items, p.cfg.radix_bits, p.cfg.cell_blocks, p.cfg.pair_blocks, p.cfg.small_max};
}
dbscan_hip_batch_plan plan_from_ints(const std::vector<int64_t>& v) {
  TORCH_CHECK(v.size() == kPlanInts, "dbscan2d_batched_enqueue: a plan has ", kPlanInts,
              " values (from dbscan2d_batched_stage)");
  dbscan_hip_batch_plan p{};
  p.total_points = v[0];
  p.num_clouds = v[1];
  p.sort_tiles = v[2];
  p.scan_tiles = v[3];
  p.all_small = static_cast<int32_t>(v[4]);
  p.cfg = dbscan_hip_config{static_cast<int32_t>(v[5]), static_cast<int32_t>(v[6]),
                            static_cast<int32_t>(v[7]), static_cast<int32_t>(v[8]),
                            static_cast<int32_t>(v[9])};
  return p;
}
void check_batch_buffers(const torch::Tensor& X, const torch::Tensor& workspace) {
  TORCH_CHECK(X.is_cuda() && workspace.is_cuda() && workspace.device() == X.device(),
              "dbscan2d_batched: X and workspace must be on one ROCm device");
  TORCH_CHECK(X.is_contiguous() && workspace.is_contiguous() &&
                  workspace.scalar_type() == torch::kUInt8,
              "dbscan2d_batched: X must be contiguous and workspace a contiguous uint8 tensor");
}
}  // namespace

std::vector<int64_t> dbscan2d_batched_stage(torch::Tensor X, std::vector<int64_t> offsets,
                                            std::vector<double> eps,
                                            std::vector<int64_t> min_samples,
                                            torch::Tensor workspace) {
  const Batch b = check_batch(X, offsets, eps, min_samples);
  check_batch_buffers(X, workspace);
  const int64_t n = X.size(0), B = static_cast<int64_t>(eps.size());
  c10::DeviceGuard guard(X.device());
  hipStream_t stream = c10::hip::getCurrentHIPStream(X.device().index()).stream();
  dbscan_hip_batch_plan plan{};
  int bad = -1;
  check_status(dbscan2d_hip_batched_stage(n, b.offsets.data(), b.eps.data(), b.ms.data(), B,
                                          workspace.data_ptr(),
                                          static_cast<size_t>(workspace.numel()), stream, &bad,
                                          config_for(n), &plan));
  return plan_to_ints(plan);
}

void dbscan2d_batched_enqueue(torch::Tensor X, torch::Tensor labels, torch::Tensor workspace,
                              std::vector<int64_t> plan, std::optional<torch::Tensor> error_word) {
  const dbscan_hip_batch_plan p = plan_from_ints(plan);
  check_batch_buffers(X, workspace);
  TORCH_CHECK(X.dim() == 2 && X.size(1) == 2 && X.scalar_type() == torch::kFloat32 &&
                  X.size(0) == p.total_points,
              "dbscan2d_batched_enqueue: X must be the staged batch's float32 (N, 2) points");
  TORCH_CHECK(labels.is_cuda() && labels.device() == X.device() && labels.is_contiguous() &&
                  labels.scalar_type() == torch::kInt32 && labels.numel() >= p.total_points,
              "dbscan2d_batched_enqueue: labels must be a contiguous int32 tensor of at least N "
              "elements on X's device");
  int32_t* err = nullptr;
  if (error_word) {
    const torch::Tensor& e = *error_word;
    TORCH_CHECK(e.scalar_type() == torch::kInt32 && e.numel() >= 1 && e.is_contiguous() &&
                    (e.is_cuda() || e.is_pinned()),
                "dbscan2d_batched_enqueue: error_word must be an int32 tensor in device or "
                "pinned host memory");
    err = e.data_ptr<int32_t>();
  }
  c10::DeviceGuard guard(X.device());
  hipStream_t stream = c10::hip::getCurrentHIPStream(X.device().index()).stream();
  check_status(dbscan2d_hip_batched_enqueue(X.data_ptr<float>(), labels.data_ptr<int32_t>(),
                                            workspace.data_ptr(),
                                            static_cast<size_t>(workspace.numel()), stream, &p,
                                            err));
}

int64_t dbscan2d_batched_workspace_bytes(int64_t total_points, int64_t num_clouds) {
  size_t bytes = 0;
  check_status(dbscan2d_hip_batched_workspace_bytes(total_points, num_clouds, &bytes));
  return static_cast<int64_t>(bytes);
}

namespace {

// `steps` timed calls of `launch(events)`; returns per-kernel means and the
// total event time.
template <class F>
std::pair<std::vector<double>, double> time_kernels(int64_t steps, F launch) {
  constexpr int K = DBSCAN_HIP_NUM_KERNELS + 1;
  std::vector<hipEvent_t> ev(static_cast<size_t>(steps) * K, nullptr);
  auto destroy = [&]() {
    for (auto& x : ev)
      if (x) (void)hipEventDestroy(x);
  };
  for (auto& x : ev) {
    if (hipEventCreate(&x) != hipSuccess) {
      destroy();
      TORCH_CHECK(false, "dbscan2d_benchmark: hipEventCreate failed");
    }
  }
  for (int64_t s = 0; s < steps; ++s) {
    const int rc = launch(ev.data() + s * K);
    if (rc != DBSCAN_HIP_OK) destroy();
    check_status(rc);
  }
  if (hipEventSynchronize(ev.back()) != hipSuccess) {
    destroy();
    TORCH_CHECK(false, "dbscan2d_benchmark: hipEventSynchronize failed");
  }
  std::vector<double> mean(DBSCAN_HIP_NUM_KERNELS, 0.0);
  for (int64_t s = 0; s < steps; ++s) {
    for (int k = 0; k < DBSCAN_HIP_NUM_KERNELS; ++k) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[s * K + k], ev[s * K + k + 1]);
      mean[k] += ms / static_cast<double>(steps);
    }
  }
  float total = 0.f;
  (void)hipEventElapsedTime(&total, ev.front(), ev.back());
  destroy();
  return {mean, static_cast<double>(total)};
}

std::vector<std::string> kernel_names() {
  std::vector<std::string> names;
  for (int k = 0; k < DBSCAN_HIP_NUM_KERNELS; ++k) names.emplace_back(dbscan2d_hip_kernel_name(k));
  return names;
}

}  // namespace

// Benchmark helper (no reference counterpart): `steps` back-to-back calls on
// one resident input, with HIP events on the launch stream before every
// kernel.  Returns (labels, per-kernel mean ms over the steps, kernel names,
// event-measured ms from the first kernel of step 0 to the end of the last).
std::tuple<torch::Tensor, std::vector<double>, std::vector<std::string>, double>
dbscan2d_benchmark(torch::Tensor X, double eps, int64_t min_samples, int64_t steps,
                   std::optional<std::vector<int64_t>> config) {
  check_input(X, eps, min_samples);
  TORCH_CHECK(X.is_cuda() && X.size(0) > 0 && steps >= 1,
              "dbscan2d_benchmark: needs a non-empty ROCm tensor and steps >= 1");
  c10::DeviceGuard guard(X.device());
  auto Xd = X.contiguous();
  const int64_t n = Xd.size(0);
  size_t ws_bytes = 0;
  check_status(dbscan2d_hip_workspace_bytes(n, &ws_bytes));
  auto opts = torch::TensorOptions().device(X.device());
  auto workspace = torch::empty({static_cast<int64_t>(ws_bytes)}, opts.dtype(torch::kUInt8));
  auto labels = torch::empty({n}, opts.dtype(torch::kInt32));
  hipStream_t stream = c10::hip::getCurrentHIPStream(X.device().index()).stream();
  dbscan_hip_config cfg{};
  const dbscan_hip_config* cp = config_for(n);
  if (config) {
    cfg = to_config(*config, n);
    cp = &cfg;
  }
  auto r = time_kernels(steps, [&](hipEvent_t* ev) {
    return dbscan2d_hip_timed(Xd.data_ptr<float>(), n, eps, min_samples,
                              labels.data_ptr<int32_t>(), workspace.data_ptr(), ws_bytes, stream,
                              ev, cp);
  });
  return {labels, r.first, kernel_names(), r.second};
}

// Benchmark helper for batches (no reference counterpart), as
// dbscan2d_benchmark: `steps` back-to-back batched calls on resident input.
std::tuple<torch::Tensor, std::vector<double>, std::vector<std::string>, double>
dbscan2d_batched_benchmark(torch::Tensor X, std::vector<int64_t> offsets, std::vector<double> eps,
                           std::vector<int64_t> min_samples, int64_t steps,
                           std::optional<std::vector<int64_t>> config) {
  const Batch b = check_batch(X, offsets, eps, min_samples);
  TORCH_CHECK(X.is_cuda() && X.size(0) > 0 && steps >= 1,
              "dbscan2d_batched_benchmark: needs a non-empty ROCm tensor and steps >= 1");
  c10::DeviceGuard guard(X.device());
  auto Xd = X.contiguous();
  const int64_t n = Xd.size(0), B = static_cast<int64_t>(eps.size());
  size_t ws_bytes = 0;
  check_status(dbscan2d_hip_batched_workspace_bytes(n, B, &ws_bytes));
  auto opts = torch::TensorOptions().device(X.device());
  auto workspace = torch::empty({static_cast<int64_t>(ws_bytes)}, opts.dtype(torch::kUInt8));
  auto labels = torch::empty({n}, opts.dtype(torch::kInt32));
  hipStream_t stream = c10::hip::getCurrentHIPStream(X.device().index()).stream();
  dbscan_hip_config cfg{};
  const dbscan_hip_config* cp = config_for(n);
  if (config) {
    cfg = to_config(*config, n);
    cp = &cfg;
  }
  auto r = time_kernels(steps, [&](hipEvent_t* ev) {
    return dbscan2d_hip_batched_timed(Xd.data_ptr<float>(), n, b.offsets.data(), b.eps.data(),
                                      b.ms.data(), B, labels.data_ptr<int32_t>(),
                                      workspace.data_ptr(), ws_bytes, stream, ev, cp);
  });
  return {labels, r.first, kernel_names(), r.second};
}

// Serving helper: the clouds of a batch copied one after another into `dst`
// (an (N, 2) float32 device tensor) at the given row offsets, by one
// asynchronous copy each on the current stream (host sources should be
// pinned for the copies to overlap).  One call for the batch keeps the
// per-request host cost to a copy submission.
void copy_clouds(torch::Tensor dst, const std::vector<torch::Tensor>& srcs,
                 const std::vector<int64_t>& rows) {
  TORCH_CHECK(dst.is_cuda() && dst.is_contiguous() && dst.scalar_type() == torch::kFloat32 &&
                  dst.dim() == 2 && dst.size(1) == 2,
              "copy_clouds: dst must be a contiguous (N, 2) float32 ROCm tensor");
  TORCH_CHECK(srcs.size() == rows.size(), "copy_clouds: one row offset per cloud");
  c10::DeviceGuard guard(dst.device());
  hipStream_t stream = c10::hip::getCurrentHIPStream(dst.device().index()).stream();
  float* base = dst.data_ptr<float>();
  for (size_t i = 0; i < srcs.size(); ++i) {
    const torch::Tensor& s = srcs[i];
    TORCH_CHECK(s.is_contiguous() && s.scalar_type() == torch::kFloat32 && s.dim() == 2 &&
                    s.size(1) == 2,
                "copy_clouds: every cloud must be a contiguous (n, 2) float32 tensor");
    TORCH_CHECK(rows[i] >= 0 && rows[i] + s.size(0) <= dst.size(0),
                "copy_clouds: a cloud does not fit at its row offset");
    if (s.size(0) == 0) continue;
    C10_HIP_CHECK(hipMemcpyAsync(base + rows[i] * 2, s.data_ptr<float>(),
                                 static_cast<size_t>(s.numel()) * sizeof(float),
                                 s.is_cuda() ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                 stream));
  }
}

// csrc/extension.cpp:65-81
PYBIND11_MODULE(_C, m) {
  m.doc() = "Native PyTorch-ROCm DBSCAN (2D, grid-hashed) on gfx950.";
  m.def("dbscan2d", &dbscan2d, pybind11::arg("X"), pybind11::arg("eps"),
        pybind11::arg("min_samples"),
        "Run DBSCAN on (N, 2) float32 tensor. Returns int32 labels (-1 = noise) on X's "
        "device. The clustering runs on the HIP device.");
  m.def("dbscan2d_profile", &dbscan2d_profile, pybind11::arg("X"), pybind11::arg("eps"),
        pybind11::arg("min_samples"),
        "Profile variant: returns (labels, {stage_name: ms}); GPU time per stage.");
  m.def("dbscan2d_hip_debug", &dbscan2d_hip_debug, pybind11::arg("X"), pybind11::arg("eps"),
        pybind11::arg("min_samples"),
        "Debug variant: returns [labels, info, perm, xs, ys, point_cell, is_core, cell_start, "
        "cell_key, cell_flags, cell_ref, uf, cell_cluster]. Not part of the public API.");
  m.def("dbscan2d_cuda_debug", &dbscan2d_cuda_debug, pybind11::arg("X"), pybind11::arg("eps"),
        pybind11::arg("min_samples"),
        "Debug variant: returns a list of per-stage intermediate tensors in the reference's "
        "order (labels, keys_sorted, perm, xs, ys, is_core, cell_start, cluster_id_per_cell, "
        "uf, sorted_labels_pre_border, sorted_labels_final). Not part of the public API.");
  m.def("check_device_errors", &check_device_errors, pybind11::arg("device") = pybind11::none(),
        "Waits for the calls whose device labels were returned without a host "
        "synchronisation and raises RuntimeError if one failed a device-side check.");
  m.def("device_error_message", [](int32_t bits) { return device_error_text(bits); },
        pybind11::arg("bits"), "Text for a device error word.");
  m.def("raise_for_device_error", [](int32_t bits) { raise_for_device_error(bits); },
        pybind11::arg("bits"), "Raises RuntimeError for a non-zero device error word.");
  m.def("inject_error_for_testing", &dbscan2d_hip_inject_error_for_testing, pybind11::arg("bits"),
        "Test hook: the next call on this thread starts with these error bits.");
  m.def("dbscan2d_configured", &dbscan2d_configured, pybind11::arg("X"), pybind11::arg("eps"),
        pybind11::arg("min_samples"), pybind11::arg("config"),
        "dbscan2d with an explicit launch configuration (sort_items, radix_bits, cell_blocks, "
        "pair_blocks).");
  m.def("set_tuning_table", &set_tuning_table, pybind11::arg("rows"),
        "Installs launch configurations by size: rows of (max_n, sort_items, radix_bits, "
        "cell_blocks, pair_blocks); an empty list restores the built-in table.");
  m.def("get_tuning_table", &get_tuning_table, "The installed tuning rows.");
  m.def("default_config", &default_config, pybind11::arg("n"),
        "The library's built-in configuration for n points.");
  m.def("dbscan2d_benchmark", &dbscan2d_benchmark, pybind11::arg("X"), pybind11::arg("eps"),
        pybind11::arg("min_samples"), pybind11::arg("steps"),
        pybind11::arg("config") = pybind11::none(),
        "Benchmark helper: (labels, per-kernel mean ms, kernel names, total event ms) over "
        "`steps` back-to-back calls. Not part of the public API.");
  m.def("dbscan2d_batched", &dbscan2d_batched, pybind11::arg("X"), pybind11::arg("offsets"),
        pybind11::arg("eps"), pybind11::arg("min_samples"),
        "Batched DBSCAN: cloud i = rows [offsets[i], offsets[i+1]) of the packed (N, 2) float32 "
        "X, clustered with eps[i], min_samples[i]. Returns the packed int32 labels, numbered "
        "from 0 within every cloud, on X's device.");
  m.def("dbscan2d_batched_out", &dbscan2d_batched_out, pybind11::arg("X"),
        pybind11::arg("offsets"), pybind11::arg("eps"), pybind11::arg("min_samples"),
        pybind11::arg("labels"), pybind11::arg("workspace"),
        pybind11::arg("error_word") = pybind11::none(),
        "Serving helper: batched DBSCAN into caller-owned device buffers on the current "
        "stream; allocates nothing. error_word (one int32, device or pinned host memory) "
        "receives the batch's device error word from its last kernel.");
  m.def("copy_clouds", &copy_clouds, pybind11::arg("dst"), pybind11::arg("srcs"),
        pybind11::arg("rows"),
        "Serving helper: clouds copied into consecutive rows of a device tensor, one "
        "asynchronous copy each on the current stream.");
  m.def("dbscan2d_batched_stage", &dbscan2d_batched_stage, pybind11::arg("X"),
        pybind11::arg("offsets"), pybind11::arg("eps"), pybind11::arg("min_samples"),
        pybind11::arg("workspace"),
        "Stages a batch's tables in `workspace` (current stream); returns the launch plan.");
  m.def("dbscan2d_batched_enqueue", &dbscan2d_batched_enqueue, pybind11::arg("X"),
        pybind11::arg("labels"), pybind11::arg("workspace"), pybind11::arg("plan"),
        pybind11::arg("error_word") = pybind11::none(),
        "Enqueues a staged batch's kernels only, on the current stream (capturable).");
  m.def("dbscan2d_batched_workspace_bytes", &dbscan2d_batched_workspace_bytes,
        pybind11::arg("total_points"), pybind11::arg("num_clouds"),
        "Workspace bytes of a batch of num_clouds clouds holding total_points points.");
  m.def("dbscan2d_batched_benchmark", &dbscan2d_batched_benchmark, pybind11::arg("X"),
        pybind11::arg("offsets"), pybind11::arg("eps"), pybind11::arg("min_samples"),
        pybind11::arg("steps"), pybind11::arg("config") = pybind11::none(),
        "Benchmark helper for batches, as dbscan2d_benchmark. Not part of the public API.");
}
