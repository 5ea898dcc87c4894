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
