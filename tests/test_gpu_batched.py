"""Batched clustering on the GPU: many clouds, each with its own eps and
min_samples, in one pipeline of launches.  Every cloud's labels must equal,
bit for bit, what the oracle (the reference CPU semantics) gives that cloud
alone, and what a plain dbscan2d call gives it, whatever else is in the
batch."""

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


def batched(dt, clouds, eps, ms):
    print(f"[gpu batch] clouds={len(clouds)} points={sum(len(c) for c in clouds)}", flush=True)
    parts = [torch.from_numpy(np.ascontiguousarray(c)).to(DEV) for c in clouds]
    out = dt.dbscan2d_batched(parts, eps, ms)
    assert len(out) == len(clouds)
    for o, c in zip(out, clouds):
        assert o.dtype == torch.int32 and o.device.type == "cuda" and o.shape == (len(c),)
    return [o.cpu().numpy() for o in out]


def _cloud(rng, n):
    kind = int(rng.integers(0, 5))
    if kind == 0:
        c = rng.uniform(-30, 30, (6, 2))
        X = np.concatenate([rng.normal(ci, 1.0, (n, 2)) for ci in c])[:n]
    elif kind == 1:
        X = rng.uniform(-50, 50, (n, 2))
    elif kind == 2:
        X = rng.integers(-40, 40, (n, 2)) * 0.25  # exact-eps ties
    elif kind == 3:
        X = np.concatenate([rng.normal(rng.uniform(-5, 5, 2), 0.02, (n, 2)),
                            rng.uniform(-6, 6, (n, 2))])[:n]
    else:
        X = rng.uniform(1e5, 1e5 + 100, (n, 2))
    return X.astype(np.float32)


def _random_batch(seed, count):
    rng = np.random.default_rng(500 + seed)
    sizes = rng.choice([0, 1, 2, 3, 64, 65, 300, 2047, 2048, 2049, 3071, 3072, 3073, 9000, 40000],
                       count)
    clouds = [_cloud(rng, int(n)) for n in sizes]
    eps = [float(e) for e in rng.choice([0.1, 0.25, 0.5, 1.0, 2.0], count)]
    ms = [int(m) for m in rng.choice([1, 2, 5, 10, 20], count)]
    return clouds, eps, ms


@pytest.mark.parametrize("seed", range(6))
def test_random_batches_match_oracle(dt, seed):
    clouds, eps, ms = _random_batch(seed, 40 + 10 * seed)
    got = batched(dt, clouds, eps, ms)
    for i, (c, e, m) in enumerate(zip(clouds, eps, ms)):
        assert np.array_equal(got[i], oracle.dbscan2d(c, e, m)), (seed, i, len(c), e, m)


def test_constructed_cases_in_one_batch(dt, golden_dir):
    """The constructed edge cases (ties at exactly eps, negative cell
    coordinates, NaN/inf, duplicates, huge and denormal eps) as one batch,
    against the reference's own labels."""
    cases = constructed_cases()
    z = np.load(os.path.join(golden_dir, "cases.npz"), allow_pickle=False)
    names = sorted(cases)
    got = batched(dt, [cases[k][0] for k in names], [cases[k][1] for k in names],
                  [cases[k][2] for k in names])
    for k, g in zip(names, got):
        assert np.array_equal(g, z[f"labels_{k}"]), k


def test_large_clouds_in_a_batch(dt, golden_dir):
    """Workloads with golden hashes (blobs10k, lattice1m) batched with small
    clouds around them keep their reference labels."""
    with open(os.path.join(golden_dir, "golden.json")) as f:
        golden = json.load(f)
    rng = np.random.default_rng(77)
    a, ea, ma = W.blobs10k()
    b, eb, mb = W.lattice1m()
    clouds = [_cloud(rng, 500), a, _cloud(rng, 3), b, _cloud(rng, 7000)]
    eps = [0.5, ea, 0.3, eb, 1.0]
    ms = [5, ma, 1, mb, 10]
    got = batched(dt, clouds, eps, ms)
    for lab, name in ((got[1], "blobs10k"), (got[3], "lattice1m")):
        h = hashlib.sha256(np.ascontiguousarray(lab, dtype=np.int32).tobytes()).hexdigest()
        assert h == golden[name]["sha256"], name
    for i in (0, 2, 4):
        assert np.array_equal(got[i], oracle.dbscan2d(clouds[i], eps[i], ms[i]))


def test_twenty_million_point_batch(dt, golden_dir):
    """blobs10m and skewed10m in one batch (past 2^24 points in all, where
    labels go straight to their input positions) keep their reference
    labels."""
    with open(os.path.join(golden_dir, "golden.json")) as f:
        golden = json.load(f)
    rng = np.random.default_rng(78)
    a, ea, ma = W.blobs10m()
    b, eb, mb = W.skewed10m()
    small = _cloud(rng, 5000)
    got = batched(dt, [a, small, b], [ea, 0.5, eb], [ma, 5, mb])
    for lab, name in ((got[0], "blobs10m"), (got[2], "skewed10m")):
        h = hashlib.sha256(np.ascontiguousarray(lab, dtype=np.int32).tobytes()).hexdigest()
        assert h == golden[name]["sha256"], name
    assert np.array_equal(got[1], oracle.dbscan2d(small, 0.5, 5))


def test_alone_and_in_the_mix(dt):
    """A cloud's labels do not depend on the other clouds of its batch nor on
    its position in it."""
    clouds, eps, ms = _random_batch(11, 30)
    mix = batched(dt, clouds, eps, ms)
    order = np.random.default_rng(5).permutation(len(clouds))
    shuffled = batched(dt, [clouds[i] for i in order], [eps[i] for i in order],
                       [ms[i] for i in order])
    for j, i in enumerate(order):
        assert np.array_equal(shuffled[j], mix[i])
    for i in (0, 7, 29):
        alone = batched(dt, [clouds[i]], [eps[i]], [ms[i]])[0]
        assert np.array_equal(alone, mix[i])
        single = dt.dbscan2d(torch.from_numpy(clouds[i]).to(DEV), eps[i], ms[i]).cpu().numpy()
        assert np.array_equal(single, mix[i])


def test_clouds2048_sample_equals_single_calls(dt):
    """The first 256 clouds of the 2,048-cloud workload (sizes capped at
    60,000): batched labels equal plain dbscan2d calls, and a sample equals the
    oracle."""
    ns, eps, ms = W.clouds2048_params()
    idx = range(256)
    clouds = [W.clouds2048_cloud(i, min(int(ns[i]), 60_000)) for i in idx]
    e = [float(eps[i]) for i in idx]
    m = [int(ms[i]) for i in idx]
    got = batched(dt, clouds, e, m)
    for i in idx:
        single = dt.dbscan2d(torch.from_numpy(clouds[i]).to(DEV), e[i], m[i]).cpu().numpy()
        assert np.array_equal(got[i], single), i
    for i in range(0, 256, 17):
        assert np.array_equal(got[i], oracle.dbscan2d(clouds[i], e[i], m[i])), i


def _digest(labels):
    return hashlib.sha256(np.ascontiguousarray(labels, dtype=np.int32).tobytes()).hexdigest()[:16]


@pytest.mark.timeout(900)
def test_clouds2048_full_batch_matches_reference_digests(dt, golden_dir):
    """All 2,048 clouds of the batched workload (135M points, per-cloud N
    log-uniform in [200, 500,000], own eps and min_samples) in one batched
    call: every cloud's labels equal the compiled reference's
    (tests/golden/streams.json, digests of its int32 labels)."""
    with open(os.path.join(golden_dir, "streams.json")) as f:
        want = json.load(f)["clouds2048"]
    ns, eps, ms = W.clouds2048_params()
    assert int(ns.sum()) == want["points"]
    offs = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    X = np.empty((int(offs[-1]), 2), np.float32)
    for i in range(len(ns)):
        X[offs[i]:offs[i + 1]] = W.clouds2048_cloud(i, int(ns[i]))
        if i % 512 == 511:
            print(f"[gpu batch] clouds2048: {i + 1} clouds generated", flush=True)
    Xd = torch.from_numpy(X).to(DEV)
    del X
    print(f"[gpu batch] clouds2048: one batched call on {int(offs[-1])} points", flush=True)
    labels = dt.dbscan2d_batched(Xd, [float(e) for e in eps], [int(m) for m in ms],
                                 offsets=offs.tolist()).cpu().numpy()
    dt.check_device_errors()
    bad = [i for i in range(len(ns)) if _digest(labels[offs[i]:offs[i + 1]]) != want["digests"][i]]
    assert not bad, f"{len(bad)} clouds differ from the reference, first {bad[:10]}"


def test_packed_input_with_empty_clouds(dt):
    rng = np.random.default_rng(21)
    sizes = [0, 1200, 0, 0, 5, 4000, 0]
    clouds = [_cloud(rng, n) for n in sizes]
    offs = np.concatenate([[0], np.cumsum(sizes)]).tolist()
    packed = torch.from_numpy(np.concatenate(clouds)).to(DEV)
    labels = dt.dbscan2d_batched(packed, [0.5] * 7, [4] * 7, offsets=offs).cpu().numpy()
    assert labels.shape == (sum(sizes),)
    for i, c in enumerate(clouds):
        assert np.array_equal(labels[offs[i]:offs[i + 1]], oracle.dbscan2d(c, 0.5, 4))
    empty = dt.dbscan2d_batched([torch.empty((0, 2), device=DEV)] * 3, 0.5, 4)
    assert [t.numel() for t in empty] == [0, 0, 0]


def test_min_samples_narrowing_per_cloud(dt):
    rng = np.random.default_rng(3)
    X = rng.normal(0, 0.5, (2000, 2)).astype(np.float32)
    got = batched(dt, [X, X, X], [0.2, 0.2, 0.2], [2**32 + 4, 2**31, 4])
    ref = oracle.dbscan2d(X, 0.2, 4)
    assert np.array_equal(got[0], ref) and np.array_equal(got[2], ref)
    assert (got[1] == -1).all()


def test_host_tensors(dt):
    clouds, eps, ms = _random_batch(2, 12)
    out = dt.dbscan2d_batched([torch.from_numpy(c) for c in clouds], eps, ms)
    for o, c, e, m in zip(out, clouds, eps, ms):
        assert o.device.type == "cpu"
        assert np.array_equal(o.numpy(), oracle.dbscan2d(c, e, m))


def test_c_abi_batched_direct(dt):
    lib = ctypes.CDLL(os.path.join(os.path.dirname(dt.__file__), "libdbscan_hip.so"))
    P = ctypes.c_void_p
    lib.dbscan2d_hip_batched.argtypes = [P, ctypes.c_int64, P, P, P, ctypes.c_int64, P, P,
                                         ctypes.c_size_t, P, ctypes.POINTER(ctypes.c_int), P]
    lib.dbscan2d_hip_batched_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int64,
                                                         ctypes.POINTER(ctypes.c_size_t)]
    clouds, eps, ms = _random_batch(4, 9)
    sizes = [len(c) for c in clouds]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    total = int(offs[-1])
    e = np.array(eps, dtype=np.float64)
    m = np.array(ms, dtype=np.int64)
    Xd = torch.from_numpy(np.concatenate(clouds)).to(DEV)
    labels = torch.full((total,), 7, dtype=torch.int32, device=DEV)
    nbytes = ctypes.c_size_t(0)
    assert lib.dbscan2d_hip_batched_workspace_bytes(total, len(clouds), ctypes.byref(nbytes)) == 0
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device=DEV)
    stream = torch.cuda.current_stream().cuda_stream
    bad = ctypes.c_int(123)

    def call(eps_arr, ws_bytes):
        return lib.dbscan2d_hip_batched(Xd.data_ptr(), total, offs.ctypes.data,
                                        eps_arr.ctypes.data, m.ctypes.data, len(clouds),
                                        labels.data_ptr(), ws.data_ptr(), ws_bytes, stream,
                                        ctypes.byref(bad), None)

    assert call(e, nbytes.value) == 0 and bad.value == -1
    torch.cuda.synchronize()
    got = labels.cpu().numpy()
    for i, c in enumerate(clouds):
        assert np.array_equal(got[offs[i]:offs[i + 1]], oracle.dbscan2d(c, eps[i], ms[i]))
    e_bad = e.copy()
    e_bad[5] = 0.0
    assert call(e_bad, nbytes.value) == 2 and bad.value == 5   # DBSCAN_HIP_ERR_EPS
    assert call(e, nbytes.value - 1) == 5                       # DBSCAN_HIP_ERR_WORKSPACE
