"""The allocation-free entry point (dbscan2d_hip_noalloc) on the GPU, through
the C-ABI: its labels equal those of dbscan2d_hip (the workspace entry point)
on three calls in a row, each on a smaller input than the one before (the
last small enough for the one-workgroup path), and the oracle's; sizes above
its limit are refused before any device work."""

import ctypes
import os
import re

import numpy as np
import pytest
import torch

import oracle
from tests import workloads as W

pytestmark = pytest.mark.gpu

DEV = "cuda"
HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "dbscan_hip.h")


@pytest.fixture(scope="module")
def lib():
    import dbscan_torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    lib = ctypes.CDLL(os.path.join(os.path.dirname(dbscan_torch.__file__), "libdbscan_hip.so"))
    P = ctypes.c_void_p
    lib.dbscan2d_hip.argtypes = [P, ctypes.c_int64, ctypes.c_double, ctypes.c_int64, P, P,
                                 ctypes.c_size_t, P]
    lib.dbscan2d_hip_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]
    lib.dbscan2d_hip_noalloc.argtypes = [P, ctypes.c_int64, ctypes.c_double, ctypes.c_int64, P, P]
    return lib


def _max_points():
    return int(re.search(r"#define DBSCAN_HIP_NOALLOC_MAX_POINTS \(1 << (\d+)\)",
                         open(HEADER).read()).group(1))


def _with_workspace(lib, Xd, eps, ms, stream):
    n = Xd.shape[0]
    nbytes = ctypes.c_size_t(0)
    assert lib.dbscan2d_hip_workspace_bytes(n, ctypes.byref(nbytes)) == 0
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device=DEV)
    out = torch.full((n,), 7, dtype=torch.int32, device=DEV)
    assert lib.dbscan2d_hip(Xd.data_ptr(), n, eps, ms, out.data_ptr(), ws.data_ptr(),
                            nbytes.value, stream) == 0
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_equals_workspace_entry_point(lib):
    stream = torch.cuda.current_stream().cuda_stream
    big = W.G(1 << _max_points(), 120, (0, 400), 1.0, 0.1, (0, 400), 31)  # the largest allowed
    small = W.G(30_000, 10, (-15, 15), 0.6, 0.1, (-18, 18), 32)
    tiny = W.G(3_000, 4, (-5, 5), 0.4, 0.1, (-6, 6), 33)  # one-workgroup path
    for X, eps, ms in ((big, 0.5, 10), (small, 0.4, 6), (tiny, 0.3, 5)):
        n = len(X)
        print(f"[gpu noalloc] n={n} eps={eps} ms={ms}", flush=True)
        Xd = torch.from_numpy(X).to(DEV)
        labels = torch.full((n,), 7, dtype=torch.int32, device=DEV)
        assert lib.dbscan2d_hip_noalloc(Xd.data_ptr(), n, eps, ms, labels.data_ptr(), stream) == 0
        torch.cuda.synchronize()
        got = labels.cpu().numpy()
        assert np.array_equal(got, _with_workspace(lib, Xd, eps, ms, stream))
        if n <= 100_000:
            assert np.array_equal(got, oracle.dbscan2d(X, eps, ms))


def test_noalloc_limits(lib):
    stream = torch.cuda.current_stream().cuda_stream
    top = 1 << _max_points()
    # refused before any device work (fake addresses, never touched)
    assert lib.dbscan2d_hip_noalloc(0x1000, top + 1, 0.5, 5, 0x2000, stream) == 10
    assert lib.dbscan2d_hip_noalloc(0x1000, 10, 0.0, 5, 0x2000, stream) == 2
    assert lib.dbscan2d_hip_noalloc(0x1000, 10, 0.5, 0, 0x2000, stream) == 3
    assert lib.dbscan2d_hip_noalloc(None, 0, 0.5, 5, None, stream) == 0
    lib.dbscan2d_hip_status_string.restype = ctypes.c_char_p
    assert b"DBSCAN_HIP_NOALLOC_MAX_POINTS" in lib.dbscan2d_hip_status_string(10)
