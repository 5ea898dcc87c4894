"""The continuous-batching server on the GPU: every served result equals the
labels of a plain dbscan2d call on the same cloud alone (and the oracle's),
whatever else shares its batch; batches overlap on the device; admission
respects the slot size and the resident cap."""

import hashlib
import json
import os

import numpy as np
import pytest
import torch

import oracle
from tests import workloads as W

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module")
def srv():
    import dbscan_torch  # noqa: F401
    from dbscan_torch import serving

    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return serving


def _requests(srv, count, seed, max_n=60_000, pinned=True):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice([0, 1, 5, 300, 2048, 3072, 3073, int(rng.integers(1, max_n))]))
        X = torch.from_numpy(W.cloud_at_density(max(n, 1), 7000 + 97 * seed + i)[:n])
        out.append(srv.Request(i, X.pin_memory() if pinned else X,
                               float(np.float32(rng.uniform(0.2, 1.0))),
                               int(rng.choice([1, 5, 10, 20]))))
    return out


def test_served_results_equal_plain_calls(srv):
    import dbscan_torch

    server = srv.Server(slots=3, slot_points=1 << 20, slot_clouds=64, cap_bytes=8e9)
    reqs = _requests(srv, 150, 1)
    arrivals = np.cumsum(np.random.default_rng(2).exponential(1 / 20000.0, len(reqs)))
    print(f"[gpu serve] {len(reqs)} requests", flush=True)
    srv.run_open_loop(server, reqs, arrivals)
    assert len(server.completed) == len(reqs)
    assert max(server.batches) > 1 and len(server.batches) < len(reqs)  # batching happened
    for r in reqs:
        plain = dbscan_torch.dbscan2d(r.X.to(DEV), r.eps, r.min_samples).cpu()
        assert r.labels.dtype == torch.int32 and torch.equal(r.labels, plain), r.rid
        assert r.done >= r.arrival
    for r in reqs[::15]:
        assert np.array_equal(r.labels.numpy(), oracle.dbscan2d(r.X.numpy(), r.eps,
                                                                r.min_samples)), r.rid
    s = srv.latency_summary(reqs)
    assert s["requests"] == len(reqs) and s["p50_ms"] <= s["p99_ms"]


def test_batches_overlap_on_the_device(srv):
    """Requests submitted while a batch runs go out in the next free slot
    without waiting for it."""
    server = srv.Server(slots=3, slot_points=4 << 20, slot_clouds=8, cap_bytes=8e9)
    big = [srv.Request(i, torch.from_numpy(W.cloud_at_density(1_500_000, 8000 + i)).pin_memory(),
                       0.5, 10) for i in range(6)]
    for r in big[:2]:
        server.submit(r)
    server.step()
    for r in big[2:]:
        server.submit(r)
    server.step()
    peak = len(server.inflight)
    server.drain()
    assert peak >= 2
    for r in big:
        ref = oracle.dbscan2d(r.X.numpy(), 0.5, 10)
        assert np.array_equal(r.labels.numpy(), ref), r.rid


def test_device_resident_requests(srv):
    server = srv.Server(slots=2, slot_points=1 << 20, slot_clouds=16, cap_bytes=4e9)
    reqs = _requests(srv, 20, 3, pinned=False)
    for r in reqs:
        r.X = r.X.to(DEV)
        server.submit(r)
    server.drain()
    for r in reqs:
        # labels stay on the device of X, as dbscan2d returns them
        assert r.labels.device == r.X.device
        assert np.array_equal(r.labels.cpu().numpy(),
                              oracle.dbscan2d(r.X.cpu().numpy(), r.eps, r.min_samples)), r.rid


def test_mixed_host_and_device_requests_in_one_batch(srv):
    """One batch holding host and device-resident requests: each request's
    labels come back where its points are, and equal the oracle's."""
    server = srv.Server(slots=1, slot_points=2 << 20, slot_clouds=32, cap_bytes=4e9)
    reqs = _requests(srv, 20, 6, pinned=True)
    for i, r in enumerate(reqs):
        if i % 2:
            r.X = r.X.to(DEV)
        server.submit(r)
    server.drain()
    assert server.batches == [sum(r.n > 0 for r in reqs)]
    for r in reqs:
        assert r.error is None and r.labels.device == r.X.device, r.rid
        assert np.array_equal(r.labels.cpu().numpy(),
                              oracle.dbscan2d(r.X.cpu().numpy(), r.eps, r.min_samples)), r.rid


def test_admission_limits(srv):
    with pytest.raises(ValueError, match="resident cap"):
        srv.Server(slots=2, slot_points=1 << 20, slot_clouds=16, cap_bytes=1e6)
    server = srv.Server(slots=1, slot_points=1000, slot_clouds=4, cap_bytes=1e9)
    with pytest.raises(ValueError, match="a slot holds 1000"):
        server.submit(srv.Request(0, torch.zeros((1001, 2)), 0.5, 5))
    with pytest.raises(RuntimeError, match="eps must be > 0"):
        server.submit(srv.Request(1, torch.zeros((10, 2)), 0.0, 5))
    # a slot takes at most slot_clouds requests and slot_points points
    reqs = [srv.Request(10 + i, torch.zeros((300, 2)), 0.5, 5) for i in range(7)]
    for r in reqs:
        server.submit(r)
    server.drain()
    assert server.batches == [3, 3, 1]
    assert all((r.labels == 0).all() for r in reqs)
    assert server.footprint_bytes(1000) <= server.arena_bytes


def _digest(labels):
    return hashlib.sha256(np.ascontiguousarray(labels, dtype=np.int32).tobytes()).hexdigest()[:16]


@pytest.mark.timeout(900)
@pytest.mark.parametrize("rate", W.SERVE_PINNED_RATES)
def test_pinned_serve_stream_matches_reference(srv, golden_dir, rate):
    """The first 300 requests of the open-loop Poisson stream at `rate` req/s
    (N mixture 70% log-U[1k,50k], 25% log-U[50k,1M], 5% log-U[1M,10M], each
    request its own cloud), served open-loop through the continuous-batching
    server: every result equals the compiled reference's labels
    (tests/golden/streams.json)."""
    with open(os.path.join(golden_dir, "streams.json")) as f:
        want = json.load(f)["serve"][str(rate)]
    arrivals, ns, eps, ms = W.serve_stream(rate, W.SERVE_PINNED_COUNT)
    reqs = []
    for i in range(W.SERVE_PINNED_COUNT):
        n_i, e_i, m_i, _ = want["requests"][i]
        assert (n_i, m_i) == (int(ns[i]), int(ms[i])) and np.float32(e_i) == eps[i]
        X = torch.from_numpy(W.serve_request_cloud(rate, i, n_i)).pin_memory()
        reqs.append(srv.Request(i, X, float(eps[i]), int(ms[i])))
        if i % 100 == 99:
            print(f"[gpu serve] rate {rate}: {i + 1} requests generated", flush=True)
    # the benchmark's size-aware configuration: requests of 1M points and more
    # in their own slots
    server = srv.Server(slots=3, slot_points=8 << 20, slot_clouds=4096, big_slots=3,
                        big_points=24 << 20, big_threshold=1 << 20)
    print(f"[gpu serve] rate {rate}: serving {len(reqs)} requests, {int(ns.sum())} points",
          flush=True)
    srv.run_open_loop(server, reqs, arrivals)
    assert len(server.completed) == len(reqs) and not server.failed
    bad = [r.rid for r in reqs if _digest(r.labels.numpy()) != want["requests"][r.rid][3]]
    assert not bad, f"{len(bad)} requests differ from the reference, first {bad[:10]}"
    s = srv.latency_summary(reqs)
    print(f"[gpu serve] rate {rate}: {s}", flush=True)
    assert s["requests"] == len(reqs)


def test_device_inputs_written_asynchronously(srv):
    """Device-resident requests whose points are still being produced on the
    caller's stream when they are submitted: the server's slot stream waits
    for them, so every result equals the oracle on the finished points."""
    server = srv.Server(slots=2, slot_points=1 << 20, slot_clouds=16, cap_bytes=4e9)
    big = torch.randn(4096, 4096, device=DEV)
    reqs = []
    for i in range(8):
        for _ in range(4):  # keep the current stream busy ahead of the points
            big = big @ big
            big = big / big.norm()
        X = (torch.rand(20_000, 2, device=DEV, generator=None) * 60.0).contiguous()
        reqs.append(srv.Request(i, X, 0.5, 5))
        server.submit(reqs[-1])
        server.step()
    server.drain()
    for r in reqs:
        assert r.error is None
        assert r.labels.is_cuda
        assert np.array_equal(r.labels.cpu().numpy(), oracle.dbscan2d(r.X.cpu().numpy(), 0.5, 5)), r.rid


def test_size_aware_admission(srv):
    """Requests of at least big_threshold points go only to the big slots;
    small requests submitted after them are launched at once on the small
    slots, and a big slot with no big request waiting takes small ones."""
    server = srv.Server(slots=2, slot_points=1 << 18, slot_clouds=8, big_slots=1,
                        big_points=1 << 20, big_threshold=100_000, cap_bytes=8e9)
    big = [srv.Request(i, torch.from_numpy(W.cloud_at_density(400_000, 8100 + i)).pin_memory(),
                       0.5, 10) for i in range(3)]
    small = [srv.Request(10 + i, torch.from_numpy(W.cloud_at_density(5_000, 8200 + i)).pin_memory(),
                         0.4, 5) for i in range(6)]
    for r in big + small:
        server.submit(r)
    assert len(server.big_queue) == 3 and len(server.queue) == 6
    server.step()
    launched = {r.rid for s in server.inflight for r in s.batch}
    # the big slot took the first two big requests (a third does not fit its
    # 1M points); the small slots took the small requests, not waiting behind
    assert {0, 1} & launched and all(r.rid in launched for r in small)
    assert all(not s.big or all(r.n >= 100_000 for r in s.batch) for s in server.inflight)
    assert all(s.big or all(r.n < 100_000 for r in s.batch) for s in server.inflight)
    server.drain()
    for r in big + small:
        assert r.error is None
        assert np.array_equal(r.labels.numpy(), oracle.dbscan2d(r.X.numpy(), r.eps, r.min_samples))
    with pytest.raises(ValueError, match="a slot holds"):
        server.submit(srv.Request(99, torch.zeros((2 << 20, 2)), 0.5, 5))
    # with nothing big waiting, the big slot serves small requests too
    more = [srv.Request(20 + i, torch.from_numpy(W.cloud_at_density(5_000, 8300 + i)).pin_memory(),
                        0.4, 5) for i in range(24)]
    for r in more:
        server.submit(r)
    server.step()
    assert any(s.big for s in server.inflight)
    server.drain()
    for r in more:
        assert np.array_equal(r.labels.numpy(), oracle.dbscan2d(r.X.numpy(), r.eps, r.min_samples))
