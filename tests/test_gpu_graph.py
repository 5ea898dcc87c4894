"""A plain dbscan2d call captured in a HIP graph (torch.cuda.graph) and
replayed.  The call only enqueues kernels on the current stream (no host
synchronisation, allocation or memset inside), so it can be captured; every
replay, with eager calls and synchronisations between the replays, must equal
the eager labels, also after the input buffer is refilled with another cloud
of the same size."""

import numpy as np
import pytest
import torch

import oracle
from tests import workloads as W

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module")
def dt():
    import dbscan_torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return dbscan_torch


def test_graph_replays_equal_eager_calls(dt):
    n = 300_000
    X = W.cloud_at_density(n, 77)
    Xd = torch.from_numpy(X).to(DEV)
    eager = dt.dbscan2d(Xd, 0.5, 10)
    assert np.array_equal(eager.cpu().numpy(), oracle.dbscan2d(X, 0.5, 10))
    side = torch.cuda.Stream()  # warm-up on a side stream before capture
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            dt.dbscan2d(Xd, 0.5, 10)
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    print(f"[gpu graph] capture n={n}", flush=True)
    with torch.cuda.graph(graph):
        gout = dt.dbscan2d(Xd, 0.5, 10)
    other = torch.from_numpy(W.cloud_at_density(50_000, 78)).to(DEV)
    for k in range(4):
        gout.fill_(-7)
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(gout, eager), k
        dt.dbscan2d(other, 0.3, 5)  # eager work between the replays
        torch.cuda.synchronize()
    X2 = W.cloud_at_density(n, 79)
    Xd.copy_(torch.from_numpy(X2))
    eager2 = dt.dbscan2d(Xd, 0.5, 10)
    for k in range(2):
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(gout, eager2), k
    assert np.array_equal(eager2.cpu().numpy(), oracle.dbscan2d(X2, 0.5, 10))


def _packed(sizes, seed0):
    clouds = [W.cloud_at_density(n, seed0 + i) for i, n in enumerate(sizes)]
    offs = [0]
    for c in clouds:
        offs.append(offs[-1] + len(c))
    return clouds, offs, np.concatenate(clouds)


@pytest.mark.parametrize("sizes", [
    # a pipeline batch (clouds above the one-workgroup limit among them)
    [3_000, 9_000, 150_000, 40_000, 700],
    # every cloud small: the persistent one-workgroup grid
    [300, 4_096, 1_200, 2_500],
], ids=["pipeline", "all_small"])
def test_batched_graph_replays_equal_eager_batches(dt, sizes):
    eps = [0.3, 0.5, 0.7, 0.4, 0.6][:len(sizes)]
    ms = [5, 10, 20, 10, 5][:len(sizes)]
    clouds, offs, packed = _packed(sizes, 500)
    Xd = torch.from_numpy(packed).to(DEV)
    eager = dt.dbscan2d_batched(Xd, eps, ms, offsets=offs)
    for i in (0, len(sizes) - 1):  # the oracle, independently of the batched path
        assert np.array_equal(eager[offs[i]:offs[i + 1]].cpu().numpy(),
                              oracle.dbscan2d(clouds[i], eps[i], ms[i])), i
    print(f"[gpu graph] batched capture sizes={sizes}", flush=True)
    g = dt.BatchedGraph(Xd, offs, eps, ms)
    other = torch.from_numpy(W.cloud_at_density(30_000, 81)).to(DEV)
    for k in range(4):
        g.labels.fill_(-7)
        out = g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), k
        g.check()
        dt.dbscan2d(other, 0.3, 5)  # eager work between the replays
        dt.dbscan2d_batched([other, other[:1000]], 0.5, 10)
        torch.cuda.synchronize()
    # other clouds of the same sizes, refilled in place
    clouds2, _, packed2 = _packed(sizes, 900)
    Xd.copy_(torch.from_numpy(packed2))
    eager2 = dt.dbscan2d_batched(Xd, eps, ms, offsets=offs)
    for k in range(2):
        out = g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager2), k
    assert np.array_equal(eager2[offs[0]:offs[1]].cpu().numpy(),
                          oracle.dbscan2d(clouds2[0], eps[0], ms[0]))
    # new eps and min_samples for the same sizes
    eps3, ms3 = [e * 0.8 for e in eps], [m + 1 for m in ms]
    eager3 = dt.dbscan2d_batched(Xd, eps3, ms3, offsets=offs)
    g.restage(eps3, ms3)
    for k in range(2):
        out = g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager3), k
    g.check()
