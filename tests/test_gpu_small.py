"""The one-workgroup path for clouds of at most 4,096 points, on the GPU.

Single calls and batches whose clouds are all that small run as one workgroup
per cloud with the cloud in LDS (k_small).  Its labels are compared bit for
bit with the oracle and with the golden vectors written by the compiled
reference, and with the multi-kernel pipeline forced on the same input
(configuration small_max = 0), over the cases that exercise each of its
stages: 0 to 8 radix passes (all points in one cell up to 64-bit keys), ky
straddling 0 (the reference-order rotation), NaN / inf / huge coordinates
(the INT32_MIN column), exact-eps ties, min_samples 1 and above n, and the
size limits 1 and 4,096.
"""

import os

import numpy as np
import pytest
import torch

import oracle
from tests import workloads as W
from tests.golden.cases import constructed_cases

pytestmark = pytest.mark.gpu

DEV = "cuda"
SMALL_MAX = 4096


@pytest.fixture(scope="module")
def dt():
    import dbscan_torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return dbscan_torch


def config(dt, n, small_max):
    c = list(dt._C.default_config(n))
    c[4] = small_max
    return c


@pytest.fixture(scope="module", autouse=True)
def small_default(dt):
    """The tests below run with every cloud of up to 4,096 points on the
    one-workgroup path (whatever the built-in threshold)."""
    saved = dt._C.get_tuning_table()
    dt._C.set_tuning_table([[2**31 - 1] + config(dt, 1000, SMALL_MAX)])
    yield
    dt._C.set_tuning_table(saved)


def both_paths(dt, X, eps, ms):
    """Labels of the one-workgroup path and of the pipeline (host arrays)."""
    print(f"[gpu small] n={len(X)} eps={eps} ms={ms}", flush=True)
    Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).to(DEV)
    small = dt._C.dbscan2d_configured(Xd, eps, ms, config(dt, len(X), SMALL_MAX)).cpu().numpy()
    pipe = dt._C.dbscan2d_configured(Xd, eps, ms, config(dt, len(X), 0)).cpu().numpy()
    return small, pipe


def cloud(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "blobs":  # straddles 0 in both coordinates
        return W.cloud_at_density(n, seed)
    if kind == "uniform":
        return rng.uniform(-20, 20, (n, 2)).astype(np.float32)
    if kind == "lattice":  # exact-eps ties at eps 0.5
        k = int(np.ceil(np.sqrt(n)))
        i, j = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
        X = np.stack([i.ravel() * 0.5, j.ravel() * 0.5], axis=1)[:n]
        return X.astype(np.float32)
    if kind == "wide":  # ~62-bit keys: 8 radix passes
        X = rng.uniform(-1e6, 1e6, (n, 2))
        X[: n // 2] = rng.normal(rng.uniform(-1e6, 1e6, 2), 0.002, (n // 2, 2))
        return X.astype(np.float32)
    if kind == "one_cell":  # every point in one cell: no radix pass
        return rng.uniform(10.0, 10.05, (n, 2)).astype(np.float32)
    if kind == "dense":  # thousands of points per cell
        return rng.normal(0.0, 0.05, (n, 2)).astype(np.float32)
    if kind == "positive":  # no straddle
        return rng.uniform(100, 140, (n, 2)).astype(np.float32)
    raise ValueError(kind)


KINDS = ["blobs", "uniform", "lattice", "wide", "one_cell", "dense", "positive"]
EPS = {"blobs": 0.5, "uniform": 0.6, "lattice": 0.5, "wide": 0.01, "one_cell": 1.0,
       "dense": 0.02, "positive": 0.7}


@pytest.mark.parametrize("n", [1, 2, 3, 64, 65, 1000, 1023, 1024, 1025, 4095, 4096])
@pytest.mark.parametrize("kind", KINDS)
def test_small_clouds_match_oracle(dt, kind, n):
    X = cloud(kind, n, 7 * n + len(kind))
    eps = EPS[kind]
    for ms in (1, 5, 20):
        want = oracle.dbscan2d(X, eps, ms)
        small, pipe = both_paths(dt, X, eps, ms)
        assert np.array_equal(small, want), (kind, n, ms)
        assert np.array_equal(pipe, want), (kind, n, ms)


@pytest.mark.parametrize("name", sorted(constructed_cases()))
def test_constructed_cases_both_paths(dt, golden_dir, name):
    """The reference's golden edge cases (NaN/inf, INT32_MIN column, ties,
    ring-2 boundaries, duplicates, ...) through both paths."""
    X, eps, ms = constructed_cases()[name]
    z = np.load(os.path.join(golden_dir, "cases.npz"), allow_pickle=False)
    want = z[f"labels_{name}"]
    if len(X) == 0:
        pytest.skip("empty input returns before any path runs")
    small, pipe = both_paths(dt, X, eps, ms)
    assert np.array_equal(small, want)
    assert np.array_equal(pipe, want)


def test_min_samples_above_n_and_huge(dt):
    X = cloud("blobs", 3000, 11)
    for ms in (3001, 2**31 - 1, 2**32 + 5):
        small, pipe = both_paths(dt, X, 0.5, ms)
        want = oracle.dbscan2d(X, 0.5, ms)
        assert np.array_equal(small, want) and np.array_equal(pipe, want)


def test_non_finite_and_huge_coordinates(dt):
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (2000, 2)).astype(np.float32)
    X[::97, 0] = np.nan
    X[::101, 1] = np.inf
    X[::103] = -np.inf
    X[::107, 0] = 3.0e38
    X[::109] = [2.0e9, -2.0e9]
    for eps, ms in ((0.3, 3), (1e-3, 1), (1e30, 4)):
        want = oracle.dbscan2d(X, eps, ms)
        small, pipe = both_paths(dt, X, eps, ms)
        assert np.array_equal(small, want), (eps, ms)
        assert np.array_equal(pipe, want), (eps, ms)


def test_repeated_small_calls_identical(dt):
    X = torch.from_numpy(cloud("blobs", 4096, 3)).to(DEV)
    first = dt.dbscan2d(X, 0.5, 10)
    for _ in range(50):
        assert torch.equal(dt.dbscan2d(X, 0.5, 10), first)
    assert np.array_equal(first.cpu().numpy(), oracle.dbscan2d(X.cpu().numpy(), 0.5, 10))


def _batch(count, seed, max_n=SMALL_MAX):
    rng = np.random.default_rng(seed)
    ns = rng.integers(0, max_n + 1, count)
    special = [max_n, 1, 0]  # the largest cloud, one point, an empty cloud
    ns[:min(3, count)] = special[:min(3, count)]
    clouds = [cloud(KINDS[i % len(KINDS)], int(k), seed * 1000 + i) if k else
              np.zeros((0, 2), np.float32) for i, k in enumerate(ns)]
    eps = [EPS[KINDS[i % len(KINDS)]] for i in range(count)]
    ms = rng.choice([1, 5, 10, 20], count).tolist()
    return clouds, eps, ms


@pytest.mark.parametrize("count", [1, 2, 300, 700])
def test_all_small_batch_matches_oracle(dt, count):
    """A batch of only small clouds runs as a persistent grid taking clouds
    from a device ticket (one workgroup when there is one cloud)."""
    clouds, eps, ms = _batch(count, count)
    print(f"[gpu small] batch of {count}", flush=True)
    out = dt.dbscan2d_batched([torch.from_numpy(c).to(DEV) for c in clouds], eps, ms)
    for i, (c, o) in enumerate(zip(clouds, out)):
        assert np.array_equal(o.cpu().numpy(), oracle.dbscan2d(c, eps[i], ms[i])), i


def test_many_small_clouds_use_the_pipeline_and_match(dt):
    """More small clouds than the persistent grid takes (six per CU): the
    batch runs the pipeline, with the same labels."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    count = 6 * cus + 17
    rng = np.random.default_rng(77)
    ns = rng.integers(1, 40, count)
    clouds = [cloud(KINDS[i % len(KINDS)], int(k), 5000 + i) for i, k in enumerate(ns)]
    eps = [EPS[KINDS[i % len(KINDS)]] for i in range(count)]
    ms = rng.choice([1, 2, 3], count).tolist()
    out = dt.dbscan2d_batched([torch.from_numpy(c).to(DEV) for c in clouds], eps, ms)
    for i, (c, o) in enumerate(zip(clouds, out)):
        assert np.array_equal(o.cpu().numpy(), oracle.dbscan2d(c, eps[i], ms[i])), i


def test_mixed_batch_uses_the_pipeline_and_matches(dt):
    clouds, eps, ms = _batch(40, 99)
    clouds.append(cloud("blobs", 6000, 5))
    eps.append(0.5)
    ms.append(10)
    out = dt.dbscan2d_batched([torch.from_numpy(c).to(DEV) for c in clouds], eps, ms)
    for i, (c, o) in enumerate(zip(clouds, out)):
        assert np.array_equal(o.cpu().numpy(), oracle.dbscan2d(c, eps[i], ms[i])), i


def test_small_batch_on_two_streams(dt):
    """Two all-small batches in flight at once (their persistent grids share
    the GPU) give each batch's own labels."""
    batches = [_batch(200, 40 + b) for b in range(2)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    want = [[oracle.dbscan2d(c, e, m) for c, e, m in zip(*b)] for b in batches]
    inputs = [[torch.from_numpy(c).to(DEV) for c in b[0]] for b in batches]
    torch.cuda.synchronize()
    for rep in range(3):
        outs = []
        for b, s in zip(range(2), streams):
            with torch.cuda.stream(s):
                outs.append(dt.dbscan2d_batched(inputs[b], batches[b][1], batches[b][2]))
        torch.cuda.synchronize()
        for b in range(2):
            for i, o in enumerate(outs[b]):
                assert np.array_equal(o.cpu().numpy(), want[b][i]), (rep, b, i)


def test_small_path_error_reporting(dt):
    """An injected error bit poisons a small call's labels and is raised, for
    a single call and for a persistent small batch."""
    X = torch.from_numpy(cloud("blobs", 2000, 1)).to(DEV)
    dt.check_device_errors()
    dt._C.inject_error_for_testing(8)
    labels = dt.dbscan2d(X, 0.5, 5)
    torch.cuda.synchronize()
    assert bool((labels == -2**31).all())
    with pytest.raises(RuntimeError, match="outside its array"):
        dt.check_device_errors()
    clouds, eps, ms = _batch(50, 3)
    dt._C.inject_error_for_testing(4)
    out = dt.dbscan2d_batched([torch.from_numpy(c).to(DEV) for c in clouds], eps, ms)
    torch.cuda.synchronize()
    assert all(bool((o == -2**31).all()) for o in out)
    with pytest.raises(RuntimeError, match="union-find operation"):
        dt.check_device_errors()
    # the next calls are clean again
    assert np.array_equal(dt.dbscan2d(X, 0.5, 5).cpu().numpy(),
                          oracle.dbscan2d(X.cpu().numpy(), 0.5, 5))
    dt.check_device_errors()


def test_small_call_in_a_hip_graph(dt):
    """A small call is one kernel launch; captured and replayed it gives the
    eager labels on every replay, also after the input is refilled."""
    X1 = torch.from_numpy(cloud("blobs", 3000, 8)).to(DEV)
    X2 = torch.from_numpy(cloud("uniform", 3000, 9)).to(DEV)
    static_x = X1.clone()
    dt.dbscan2d(static_x, 0.5, 5)  # warm-up outside the capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        static_out = dt.dbscan2d(static_x, 0.5, 5)
    for src in (X1, X2, X1):
        static_x.copy_(src)
        eager = dt.dbscan2d(src, 0.5, 5)
        torch.cuda.synchronize()
        for _ in range(3):
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(static_out, eager)
        assert np.array_equal(eager.cpu().numpy(), oracle.dbscan2d(src.cpu().numpy(), 0.5, 5))
