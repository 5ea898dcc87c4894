"""The continuous-batching server on the GPU: every served result equals the
labels of a plain dbscan2d call on the same cloud alone (and the oracle's),
whatever else shares its batch; batches overlap on the device; admission
respects the slot size and the resident cap."""

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
        assert np.array_equal(r.labels.numpy(),
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
