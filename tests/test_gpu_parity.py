"""Parity of the HIP path with the reference CPU path, on the GPU.

Bit-exact label equality (raw cluster ids, not partitions) against:
  * golden vectors written by the compiled reference (tests/golden/);
  * the C oracle, on the same seeded inputs, at every workload size;
plus size-independent properties at full size (permutation equivariance,
run-to-run identity) and stage-wise intermediates against the oracle.
"""

import ctypes
import hashlib
import json
import os

import numpy as np
import pytest
import torch

import oracle
from tests import workloads as W
from tests.golden.cases import constructed_cases

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module")
def dt():
    import dbscan_torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return dbscan_torch


@pytest.fixture(scope="module")
def golden(golden_dir):
    with open(os.path.join(golden_dir, "golden.json")) as f:
        return json.load(f)


def sha(labels):
    return hashlib.sha256(np.ascontiguousarray(labels, dtype=np.int32).tobytes()).hexdigest()


def gpu_labels(dt, X, eps, ms):
    print(f"[gpu] n={len(X)} eps={eps} ms={ms}", flush=True)
    out = dt.dbscan2d(torch.from_numpy(np.ascontiguousarray(X)).to(DEV), eps, ms)
    assert out.device.type == "cuda" and out.dtype == torch.int32
    return out.cpu().numpy()


def error_flag(dt, X, eps, ms):
    dbg = dt._C.dbscan2d_hip_debug(torch.from_numpy(X).to(DEV), eps, ms)
    return int(dbg[1][6].item())


@pytest.mark.parametrize("name", sorted(constructed_cases()))
def test_constructed_cases(dt, golden_dir, name):
    X, eps, ms = constructed_cases()[name]
    z = np.load(os.path.join(golden_dir, "cases.npz"), allow_pickle=False)
    assert np.array_equal(gpu_labels(dt, X, eps, ms), z[f"labels_{name}"])


def _random_case(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.choice([1, 2, 3, 63, 64, 65, 4095, 4096, 4097, 8191, 20000, 70000]))
    kind = seed % 5
    if kind == 0:
        c = rng.uniform(-30, 30, (8, 2))
        X = np.concatenate([rng.normal(ci, 1.0, (n, 2)) for ci in c])[:n]
    elif kind == 1:
        X = rng.uniform(-50, 50, (n, 2))
    elif kind == 2:
        X = rng.integers(-40, 40, (n, 2)) * 0.25
    elif kind == 3:
        X = np.concatenate([rng.normal(rng.uniform(-5, 5, 2), 0.02, (n, 2)),
                            rng.uniform(-6, 6, (n, 2))])[:n]
    else:
        X = rng.uniform(1e5, 1e5 + 100, (n, 2))
    eps = float(rng.choice([0.1, 0.25, 0.5, 1.0, 2.0]))
    ms = int(rng.choice([1, 2, 5, 10, 20]))
    return X.astype(np.float32), eps, ms


@pytest.mark.parametrize("seed", range(40))
def test_random_inputs_match_oracle(dt, seed):
    X, eps, ms = _random_case(seed)
    assert np.array_equal(gpu_labels(dt, X, eps, ms), oracle.dbscan2d(X, eps, ms))


@pytest.mark.parametrize("ms", [65, 100, 300])
def test_min_samples_above_64(dt, ms):
    """min_samples above 64: cells of 65 to ms - 1 points are sparse and
    larger than a lane's scan, so core marking counts their points with the
    whole wave and border assignment searches them with the wave; labels
    equal the oracle's on the pipeline (63,000 points), on the one-workgroup
    path (one blob of 3,000 points) and in a batch of both."""
    rng = np.random.default_rng(100 + ms)
    blobs = np.concatenate([rng.normal(c, 0.3, (3000, 2)) for c in rng.uniform(0, 30, (20, 2))])
    X = np.concatenate([blobs, rng.uniform(-2, 32, (3000, 2))]).astype(np.float32)
    X = X[rng.permutation(len(X))]
    small = blobs[:3000].astype(np.float32)  # one blob: dense cells in a small cloud
    want, want_small = oracle.dbscan2d(X, 0.5, ms), oracle.dbscan2d(small, 0.5, ms)
    assert (want >= 0).any() and (want == -1).any()
    assert np.array_equal(gpu_labels(dt, X, 0.5, ms), want)
    assert np.array_equal(gpu_labels(dt, small, 0.5, ms), want_small)
    out = dt.dbscan2d_batched([torch.from_numpy(X).to(DEV), torch.from_numpy(small).to(DEV)],
                              [0.5, 0.5], [ms, ms])
    assert np.array_equal(out[0].cpu().numpy(), want)
    assert np.array_equal(out[1].cpu().numpy(), want_small)


def test_blobs10k_golden(dt, golden_dir, golden):
    X, eps, ms = W.blobs10k()
    labels = gpu_labels(dt, X, eps, ms)
    assert np.array_equal(labels, np.load(os.path.join(golden_dir, "blobs10k_labels.npy")))
    assert sha(labels) == golden["blobs10k"]["sha256"]


@pytest.mark.parametrize("name", ["lattice1m", "blobs10m", "skewed10m"])
def test_large_configs(dt, golden, name):
    X, eps, ms = W.WORKLOADS[name]()
    labels = gpu_labels(dt, X, eps, ms)
    g = golden[name]
    assert int(labels.max()) + 1 == g["n_clusters"]
    assert int((labels == -1).sum()) == g["n_noise"]
    assert sha(labels) == g["sha256"]
    assert np.array_equal(labels, oracle.dbscan2d(X, eps, ms))
    assert error_flag(dt, X, eps, ms) == 0


def test_twenty_million_points(dt):
    """Past 2^24 points the labels go straight to their input positions
    instead of through the bucketed two-pass placement: blobs10m and a copy
    shifted by 4096 in x and y, against the oracle."""
    X, eps, ms = W.blobs10m()
    X2 = np.concatenate([X, X + np.float32(4096.0)])
    labels = gpu_labels(dt, X2, eps, ms)
    assert np.array_equal(labels, oracle.dbscan2d(X2, eps, ms))


@pytest.mark.parametrize("clusters", [131_070, 131_071, 300_000])
def test_label_pair_forms(dt, clusters):
    """Isolated points with min_samples 1 are one cluster each.  Below
    131,071 clusters a label pair travels as one packed word, from there on as
    two words; both sides of the switch against the oracle."""
    side = int(np.ceil(np.sqrt(clusters)))
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side)), -1).reshape(-1, 2)[:clusters]
    X = g[np.random.default_rng(clusters).permutation(clusters)].astype(np.float32)
    labels = gpu_labels(dt, X, 0.5, 1)
    assert int(labels.max()) + 1 == clusters
    assert np.array_equal(labels, oracle.dbscan2d(X, 0.5, 1))


@pytest.mark.parametrize("clusters", [100_000, 300_000])
def test_dense_cell_long_runs(dt, clusters):
    """A cell of 200,000 points sorts in input order, so the label pass's
    tiles inside it hold runs of hundreds of pairs for a few buckets; those
    buckets are unpermuted in 64-pair chunks shared by every wave.  Isolated
    points (one cluster each, both pair forms) plus the dense cell, shuffled,
    against the oracle."""
    side = int(np.ceil(np.sqrt(clusters)))
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side)), -1).reshape(-1, 2)[:clusters]
    rng = np.random.default_rng(clusters + 1)
    dense = rng.uniform(-10.1, -10.0, (200_000, 2))
    X = np.concatenate([g, dense]).astype(np.float32)
    X = X[rng.permutation(len(X))]
    labels = gpu_labels(dt, X, 0.5, 1)
    assert int(labels.max()) + 1 == clusters + 1
    assert np.array_equal(labels, oracle.dbscan2d(X, 0.5, 1))


def test_twelve_million_points_wide_buckets(dt):
    """Between 10.5M and 2^24 points the label buckets are wider than 20,480
    positions and the unpermute runs its 1,024-thread, 128 KB workgroups;
    a cell of 1.5M points gives some buckets long runs (the chunked path of
    that instantiation).  blobs10m, the dense cell and 0.5M shifted points,
    shuffled, against the oracle."""
    X, eps, ms = W.blobs10m()
    rng = np.random.default_rng(77)
    dense = rng.uniform(3000.0, 3000.2, (1_500_000, 2)).astype(np.float32)
    X2 = np.concatenate([X, dense, X[:500_000] + np.float32(2000.0)])
    X2 = X2[rng.permutation(len(X2))]
    labels = gpu_labels(dt, X2, eps, ms)
    assert np.array_equal(labels, oracle.dbscan2d(X2, eps, ms))


def test_back_to_back_calls_reuse_the_workspace(dt):
    """Calls that reuse a workspace must not see the previous call's look-back
    words (a stale inclusive prefix once gave the last scan tile prefix 0):
    200 calls on one input, interleaved with calls on a larger one, all equal
    to the oracle."""
    side = 363
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side)), -1).reshape(-1, 2)[:131_070]
    X = g[np.random.default_rng(5).permutation(len(g))].astype(np.float32)
    big = np.concatenate([X, X + np.float32(1000.0), X + np.float32(2000.0)])
    Xd, bigd = torch.from_numpy(X).to(DEV), torch.from_numpy(big).to(DEV)
    ref = torch.from_numpy(oracle.dbscan2d(X, 0.5, 1)).to(DEV)
    print(f"[gpu] n={len(X)} x200, n={len(big)} x20", flush=True)
    for it in range(200):
        if it % 10 == 0:
            dt.dbscan2d(bigd, 0.5, 1)
        assert torch.equal(dt.dbscan2d(Xd, 0.5, 1), ref), it


def test_wide_keys_take_the_tail_passes(dt):
    """Cell keys wider than 32 bits (x and y spread over 2e6 with eps 0.01:
    about 29 bits each) need more than four radix passes, run by the tail
    launch whose workgroups claim tiles by ticket; single and batched calls
    against the oracle."""
    rng = np.random.default_rng(41)
    far = rng.uniform(-1e6, 1e6, (40_000, 2))
    blobs = np.concatenate([rng.normal(c, 0.004, (500, 2)) for c in rng.uniform(-1e6, 1e6, (20, 2))])
    X = np.concatenate([far, blobs]).astype(np.float32)
    X = X[rng.permutation(len(X))]
    dbg = dt._C.dbscan2d_hip_debug(torch.from_numpy(X).to(DEV), 0.01, 5)
    assert int(dbg[1][5].item()) > 4  # radix passes
    ref = oracle.dbscan2d(X, 0.01, 5)
    assert np.array_equal(gpu_labels(dt, X, 0.01, 5), ref)
    small = X[:3000].copy()
    out = dt.dbscan2d_batched([torch.from_numpy(small).to(DEV), torch.from_numpy(X).to(DEV)],
                              [0.5, 0.01], [3, 5])
    assert np.array_equal(out[1].cpu().numpy(), ref)
    assert np.array_equal(out[0].cpu().numpy(), oracle.dbscan2d(small, 0.5, 3))


def _key_bits(X, eps):
    """Bits of the compressed cell key of X: bit widths of the kx and ky
    ranges at cell side eps / sqrt(2), as the device computes them."""
    inv = np.float32(1.0) / np.float32(np.float32(eps) / np.float32(np.sqrt(2.0)))
    k = np.floor(X.astype(np.float32) * inv).astype(np.int64)
    return sum(int(k[:, d].max() - k[:, d].min()).bit_length() for d in (0, 1))


@pytest.mark.parametrize("bits", [18, 27])
def test_nine_bit_digits(dt, bits):
    """Radix digits of 9 bits: an 18-bit key sorts in 2 passes and a 27-bit key
    in 3 when radix_bits is 9 (6-bit digits at radix_bits 8, 7-bit at 7), with
    every tile shape, against the oracle."""
    if bits == 18:
        X = W.G(300_000, 60, (0, 140), 0.7, 0.05, (0, 140), 12)
    else:
        X = W.G(1_000_000, 300, (0, 4000), 1.0, 0.1, (0, 4000), 11)
        X[:, 1] *= np.float32(0.5)
    X = np.clip(X, 0, None).astype(np.float32)
    assert _key_bits(X, 0.5) == bits
    ref = oracle.dbscan2d(X, 0.5, 10)
    Xd = torch.from_numpy(X).to(DEV)
    for cfg in ((12, 9, 8192, 8192), (8, 9, 2048, 16384), (16, 9, 64, 64), (12, 7, 8192, 8192)):
        print(f"[gpu] {bits}-bit key, config {cfg}", flush=True)
        got = dt._C.dbscan2d_configured(Xd, 0.5, 10, list(cfg))
        assert np.array_equal(got.cpu().numpy(), ref), cfg


def test_permutation_equivariance_lattice1m(dt):
    X, eps, ms = W.lattice1m()
    base = gpu_labels(dt, X, eps, ms)
    perm = np.random.default_rng(9).permutation(len(X))
    assert np.array_equal(gpu_labels(dt, X[perm], eps, ms), base[perm])


def test_repeated_runs_identical_skewed10m(dt):
    X, eps, ms = W.skewed10m()
    Xd = torch.from_numpy(X).to(DEV)
    first = dt.dbscan2d(Xd, eps, ms)
    for _ in range(2):
        assert torch.equal(first, dt.dbscan2d(Xd, eps, ms))


def test_stage_intermediates_match_oracle(dt):
    """Core flags, cell count, cell core status and cluster count agree with
    the oracle's stages (compared per original point index)."""
    X, eps, ms = W.lattice1m()
    ref = oracle.dbscan2d_debug(X, eps, ms)
    dbg = [t.cpu().numpy() for t in dt._C.dbscan2d_hip_debug(torch.from_numpy(X).to(DEV), eps, ms)]
    labels, info, perm, xs, ys, point_cell, is_core = dbg[:7]
    cell_start, cell_key, cell_flags, cell_ref, uf, cell_cluster = dbg[7:]
    n = len(X)
    assert info[4] == ref["num_cells"] and info[6] == 0
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.array_equal(xs, X[perm, 0]) and np.array_equal(ys, X[perm, 1])
    core_gpu = np.zeros(n, np.uint8)
    core_gpu[perm] = is_core
    core_ref = np.zeros(n, np.uint8)
    core_ref[ref["perm"]] = ref["is_core"]
    assert np.array_equal(core_gpu, core_ref)
    # cell sizes, in reference cell order
    sizes_gpu = np.diff(cell_start)
    order = np.empty_like(cell_ref)
    order[cell_ref] = np.arange(len(cell_ref))
    assert np.array_equal(sizes_gpu[order], np.diff(ref["cell_start"]))
    # union-find roots are minimum reference indices, one per component
    assert np.array_equal(uf[uf], uf) and (uf <= np.arange(len(uf))).all()
    status_ref = ref["cell_status"]
    has_core_gpu = (cell_flags & 1)[order]
    assert np.array_equal(has_core_gpu.astype(bool), status_ref != 0)
    assert np.array_equal(cell_cluster[order], ref["cell_cluster"])
    assert int(labels.max()) + 1 == ref["n_clusters"]


def test_reference_debug_binding_matches_oracle(dt):
    """_C.dbscan2d_cuda_debug returns the reference's 11 tensors in its order
    (csrc/cuda/dbscan_cuda.cu:772-800), in the reference CPU path's point and
    cell order; each equals the oracle's stage on lattice1m (keys at exact-eps
    lattice ties, ky straddling no sign, 182k cells)."""
    X, eps, ms = W.lattice1m()
    ref = oracle.dbscan2d_debug(X, eps, ms)
    out = [t.cpu().numpy() for t in dt._C.dbscan2d_cuda_debug(torch.from_numpy(X).to(DEV), eps, ms)]
    assert len(out) == 11
    labels, keys, perm, xs, ys, core, cstart, ccl, uf, pre, fin = out
    assert keys.dtype == np.int64 and perm.dtype == np.int32 and core.dtype == np.uint8
    assert np.array_equal(labels, ref["labels"])
    sizes = np.diff(ref["cell_start"])
    key_ref = (ref["cell_kx"].astype(np.int64) << 32) | (ref["cell_ky"].astype(np.int64) & 0xFFFFFFFF)
    assert np.array_equal(keys, np.repeat(key_ref, sizes))
    assert (np.diff(keys) >= 0).all()  # signed ascending
    assert np.array_equal(perm, ref["perm"])
    assert np.array_equal(xs, X[ref["perm"], 0]) and np.array_equal(ys, X[ref["perm"], 1])
    assert np.array_equal(core, ref["is_core"])
    assert np.array_equal(cstart, ref["cell_start"])
    assert np.array_equal(ccl, ref["cell_cluster"])
    assert np.array_equal(uf, ref["cell_root"])
    core_lab = np.repeat(ref["cell_cluster"], sizes)
    assert np.array_equal(pre, np.where(ref["is_core"] == 1, core_lab, -1))
    assert np.array_equal(fin, ref["labels"][ref["perm"]])


def test_reference_debug_binding_negative_keys(dt):
    """Keys of cells with kx < 0 and ky < 0 (signed order across zero)."""
    X = W.G(40_000, 10, (-15, 15), 0.7, 0.1, (-18, 18), 5)
    ref = oracle.dbscan2d_debug(X, 0.5, 5)
    out = [t.cpu().numpy() for t in dt._C.dbscan2d_cuda_debug(torch.from_numpy(X).to(DEV), 0.5, 5)]
    sizes = np.diff(ref["cell_start"])
    key_ref = (ref["cell_kx"].astype(np.int64) << 32) | (ref["cell_ky"].astype(np.int64) & 0xFFFFFFFF)
    assert (ref["cell_kx"] < 0).any() and (ref["cell_ky"] < 0).any()
    assert np.array_equal(out[1], np.repeat(key_ref, sizes))
    assert np.array_equal(out[2], ref["perm"]) and np.array_equal(out[6], ref["cell_start"])
    assert np.array_equal(out[10], ref["labels"][ref["perm"]])


def test_host_tensor_in_host_tensor_out(dt):
    X, eps, ms = W.blobs10k()
    out = dt.dbscan2d(torch.from_numpy(X), eps, ms)
    assert out.device.type == "cpu" and out.dtype == torch.int32
    assert np.array_equal(out.numpy(), oracle.dbscan2d(X, eps, ms))


def test_profile_variant(dt):
    X, eps, ms = W.blobs10k()
    labels, stages = dt._C.dbscan2d_profile(torch.from_numpy(X).to(DEV), eps, ms)
    assert np.array_equal(labels.cpu().numpy(), oracle.dbscan2d(X, eps, ms))
    assert len(stages) == 10 and all(v >= 0.0 for v in stages.values())
    assert stages["sort_gather"] > 0.0


def test_c_abi_direct(dt):
    """Calls libdbscan_hip.so through ctypes with torch-allocated device
    memory and torch's current stream, as a non-torch caller would."""
    lib = ctypes.CDLL(os.path.join(os.path.dirname(dt.__file__), "libdbscan_hip.so"))
    P = ctypes.c_void_p
    lib.dbscan2d_hip.argtypes = [P, ctypes.c_int64, ctypes.c_double, ctypes.c_int64, P, P,
                                 ctypes.c_size_t, P]
    lib.dbscan2d_hip_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]
    X, eps, ms = W.blobs10k()
    n = len(X)
    Xd = torch.from_numpy(X).to(DEV)
    labels = torch.full((n,), 7, dtype=torch.int32, device=DEV)
    nbytes = ctypes.c_size_t(0)
    assert lib.dbscan2d_hip_workspace_bytes(n, ctypes.byref(nbytes)) == 0
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device=DEV)
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.dbscan2d_hip(Xd.data_ptr(), n, eps, ms, labels.data_ptr(), ws.data_ptr(),
                          nbytes.value, stream)
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(labels.cpu().numpy(), oracle.dbscan2d(X, eps, ms))
    # too small a workspace is refused before any launch
    assert lib.dbscan2d_hip(Xd.data_ptr(), n, eps, ms, labels.data_ptr(), ws.data_ptr(),
                            nbytes.value - 1, stream) == 5


def test_min_samples_narrowing(dt):
    rng = np.random.default_rng(3)
    X = rng.normal(0, 0.5, (2000, 2)).astype(np.float32)
    assert np.array_equal(gpu_labels(dt, X, 0.2, 2**32 + 4), oracle.dbscan2d(X, 0.2, 4))
    assert (gpu_labels(dt, X, 0.2, 2**31) == -1).all()


def test_one_native_hip_runtime(dt):
    """The kernel library shares torch's HIP runtime (one libamdhip64 mapped)."""
    maps = open("/proc/self/maps").read().splitlines()
    hip = {l.split()[-1] for l in maps if "libamdhip64" in l}
    assert len(hip) == 1, hip
    assert any("libdbscan_hip.so" in l for l in maps)


def test_eight_byte_aligned_input(dt):
    """A view starting one row into its storage is only 8-byte aligned; the
    bounds and histogram kernels then read it one point at a time."""
    X, eps, ms = W.blobs10k()
    big = torch.zeros((len(X) + 1, 2), dtype=torch.float32, device=DEV)
    big[1:] = torch.from_numpy(X).to(DEV)
    view = big[1:]
    assert view.is_contiguous() and view.data_ptr() % 16 == 8
    assert np.array_equal(dt.dbscan2d(view, eps, ms).cpu().numpy(), oracle.dbscan2d(X, eps, ms))


_CONFIGS = [(i, b, cb, pb) for i in (8, 12, 16) for b in (6, 7, 8, 9)
            for cb, pb in ((64, 64), (65536, 65536))]


@pytest.fixture(scope="module")
def lattice_ref():
    X, eps, ms = W.lattice1m()
    return X, eps, ms, oracle.dbscan2d(X, eps, ms)


@pytest.mark.parametrize("cfg", _CONFIGS, ids=[f"i{c[0]}b{c[1]}g{c[2]}" for c in _CONFIGS])
def test_every_launch_configuration_gives_the_reference_labels(dt, lattice_ref, cfg):
    X, eps, ms, ref = lattice_ref
    print(f"[gpu] config {cfg}", flush=True)
    got = dt._C.dbscan2d_configured(torch.from_numpy(X).to(DEV), eps, ms, list(cfg))
    assert np.array_equal(got.cpu().numpy(), ref)


def test_batched_launch_configurations(dt):
    rng = np.random.default_rng(8)
    clouds = [W.cloud_at_density(int(n), 300 + i) for i, n in
              enumerate(rng.integers(1, 40_000, 24))]
    eps = [float(np.float32(e)) for e in rng.uniform(0.2, 1.0, 24)]
    ms = [int(m) for m in rng.choice([5, 10, 20], 24)]
    offs = np.concatenate([[0], np.cumsum([len(c) for c in clouds])]).tolist()
    packed = torch.from_numpy(np.concatenate(clouds)).to(DEV)
    refs = [oracle.dbscan2d(c, e, m) for c, e, m in zip(clouds, eps, ms)]
    for cfg in ((8, 6, 64, 64), (16, 8, 65536, 65536), (12, 7, 1024, 4096)):
        labels = dt._C.dbscan2d_batched_benchmark(packed, offs, eps, ms, 1, list(cfg))[0]
        labels = labels.cpu().numpy()
        for i, r in enumerate(refs):
            assert np.array_equal(labels[offs[i]:offs[i + 1]], r), (cfg, i)


def test_autotune_bucket_and_table(dt, tmp_path):
    from dbscan_torch import autotune

    saved = dt._C.get_tuning_table()
    try:
        table = autotune.autotune(buckets=(20_000, 200_000), steps=1,
                                  save=str(tmp_path / "t.json"))
        assert [r["max_n"] for r in table["rows"]] == [63245, 2**31 - 1]
        for b in table["buckets"]:
            assert b["ms"] <= b["default_ms"] and len(b["tried"]) > 10
        assert autotune.load(str(tmp_path / "t.json"))
        X, eps, ms = W.blobs10k()
        got = dt.dbscan2d(torch.from_numpy(X).to(DEV), eps, ms).cpu().numpy()
        assert np.array_equal(got, oracle.dbscan2d(X, eps, ms))
    finally:
        dt._C.set_tuning_table(saved)
