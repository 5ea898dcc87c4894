"""Device-side failure reporting and the radix sort's tail passes on the GPU.

Failures are never forced inside a kernel here: the library's test hook starts
the next call with chosen bits in its error word, so every kernel runs
normally while the reporting path (poisoned labels, the error word through
the C-ABI, RuntimeError from the binding, failed requests in the server) is
exercised.  The tail-pass test runs two batches of wide-key clouds on two
streams at once, the case in which tail workgroups of one launch may not all
be resident."""

import ctypes
import os

import numpy as np
import pytest
import torch

import oracle
from tests import workloads as W

pytestmark = pytest.mark.gpu

DEV = "cuda"
POISON = np.iinfo(np.int32).min


@pytest.fixture(scope="module")
def dt():
    import dbscan_torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    dbscan_torch.check_device_errors()
    return dbscan_torch


def _blobs(seed, n=20_000):
    return W.G(n, 12, (-20, 20), 0.6, 0.1, (-24, 24), seed)


def test_injected_error_poisons_labels_and_raises(dt):
    X = _blobs(1)
    Xd = torch.from_numpy(X).to(DEV)
    ref = oracle.dbscan2d(X, 0.4, 6)
    print("[gpu errors] injected bit 2, device input", flush=True)
    dt._C.inject_error_for_testing(2)
    bad = dt.dbscan2d(Xd, 0.4, 6)  # returns without waiting
    torch.cuda.synchronize()
    assert (bad.cpu().numpy() == POISON).all()
    with pytest.raises(RuntimeError, match=r"error bits 0x2: a scan look-back wait timed out"):
        dt.check_device_errors()
    dt.check_device_errors()  # reported once
    # the next call raises for a failed earlier one that has finished
    dt._C.inject_error_for_testing(4)
    dt.dbscan2d(Xd, 0.4, 6)
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match=r"an earlier call's device-side check failed"):
        dt.dbscan2d(Xd, 0.4, 6)
    # clean calls afterwards
    assert np.array_equal(dt.dbscan2d(Xd, 0.4, 6).cpu().numpy(), ref)
    dt.check_device_errors()


def test_injected_error_host_input_raises_at_once(dt):
    X = _blobs(2)
    dt._C.inject_error_for_testing(8)
    with pytest.raises(RuntimeError, match="a computed index fell outside its array"):
        dt.dbscan2d(torch.from_numpy(X), 0.4, 6)
    dt._C.inject_error_for_testing(1)
    with pytest.raises(RuntimeError, match="radix-sort look-back"):
        dt.DBSCAN(eps=0.4, min_samples=6).fit(X)
    assert np.array_equal(dt.dbscan2d(torch.from_numpy(X), 0.4, 6).numpy(),
                          oracle.dbscan2d(X, 0.4, 6))


def test_injected_error_batched(dt):
    clouds = [_blobs(3 + i, 5000) for i in range(3)]
    dt._C.inject_error_for_testing(4)
    out = dt.dbscan2d_batched([torch.from_numpy(c).to(DEV) for c in clouds], 0.4, 6)
    torch.cuda.synchronize()
    assert all((o.cpu().numpy() == POISON).all() for o in out)
    with pytest.raises(RuntimeError, match="union-find operation reached its step limit"):
        dt.check_device_errors()
    dt._C.inject_error_for_testing(4)
    with pytest.raises(RuntimeError, match="union-find"):
        dt.dbscan2d_batched([torch.from_numpy(c) for c in clouds], 0.4, 6)
    out = dt.dbscan2d_batched([torch.from_numpy(c).to(DEV) for c in clouds], 0.4, 6)
    for o, c in zip(out, clouds):
        assert np.array_equal(o.cpu().numpy(), oracle.dbscan2d(c, 0.4, 6))
    dt.check_device_errors()


def test_error_word_through_the_c_abi(dt):
    lib = ctypes.CDLL(os.path.join(os.path.dirname(dt.__file__), "libdbscan_hip.so"))
    P = ctypes.c_void_p
    lib.dbscan2d_hip.argtypes = [P, ctypes.c_int64, ctypes.c_double, ctypes.c_int64, P, P,
                                 ctypes.c_size_t, P]
    lib.dbscan2d_hip_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]
    lib.dbscan2d_hip_error_async.argtypes = [P, P, P]
    lib.dbscan2d_hip_check.argtypes = [P, P, P]
    lib.dbscan2d_hip_inject_error_for_testing.argtypes = [ctypes.c_int32]
    lib.dbscan2d_hip_inject_error_for_testing.restype = None
    X = _blobs(7)
    n = len(X)
    Xd = torch.from_numpy(X).to(DEV)
    labels = torch.empty(n, dtype=torch.int32, device=DEV)
    nbytes = ctypes.c_size_t(0)
    assert lib.dbscan2d_hip_workspace_bytes(n, ctypes.byref(nbytes)) == 0
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device=DEV)
    stream = torch.cuda.current_stream().cuda_stream
    word_d = torch.full((1,), -7, dtype=torch.int32, device=DEV)
    bits = ctypes.c_int32(-1)
    for inject in (0, 2 | 8):
        print(f"[gpu errors] C-ABI call, injected {inject}", flush=True)
        lib.dbscan2d_hip_inject_error_for_testing(inject)
        assert lib.dbscan2d_hip(Xd.data_ptr(), n, 0.4, 6, labels.data_ptr(), ws.data_ptr(),
                                nbytes.value, stream) == 0
        assert lib.dbscan2d_hip_error_async(ws.data_ptr(), word_d.data_ptr(), stream) == 0
        rc = lib.dbscan2d_hip_check(ws.data_ptr(), stream, ctypes.byref(bits))
        assert rc == (9 if inject else 0) and bits.value == inject
        assert int(word_d.item()) == inject
        got = labels.cpu().numpy()
        if inject:
            assert (got == POISON).all()
        else:
            assert np.array_equal(got, oracle.dbscan2d(X, 0.4, 6))
    # the checked entry point: the last kernel writes the word (pinned host memory)
    lib.dbscan2d_hip_checked.argtypes = [P, ctypes.c_int64, ctypes.c_double, ctypes.c_int64, P,
                                         P, ctypes.c_size_t, P, P, P]
    word_h = torch.full((1,), -7, dtype=torch.int32).pin_memory()
    for inject in (4, 0):
        lib.dbscan2d_hip_inject_error_for_testing(inject)
        assert lib.dbscan2d_hip_checked(Xd.data_ptr(), n, 0.4, 6, labels.data_ptr(),
                                        ws.data_ptr(), nbytes.value, stream, None,
                                        word_h.data_ptr()) == 0
        torch.cuda.synchronize()
        assert int(word_h[0]) == inject
        assert ((labels.cpu().numpy() == POISON).all() if inject else
                np.array_equal(labels.cpu().numpy(), oracle.dbscan2d(X, 0.4, 6)))


def test_server_marks_failed_requests(dt):
    from dbscan_torch.serving import Request, Server

    server = Server(slots=2, slot_points=1 << 18, slot_clouds=8, cap_bytes=2e9)
    clouds = [_blobs(20 + i, 3000) for i in range(4)]
    reqs = [Request(i, torch.from_numpy(c).pin_memory(), 0.4, 6) for i, c in enumerate(clouds)]
    for r in reqs[:2]:
        server.submit(r)
    dt._C.inject_error_for_testing(8)
    server.step()  # launches the first batch with the injected word
    server.drain()
    for r in reqs[2:]:
        server.submit(r)
    server.drain()
    assert sorted(server.failed) == [0, 1]
    for r in reqs[:2]:
        assert r.error and "outside its array" in r.error and (r.labels.numpy() == POISON).all()
    for r, c in zip(reqs[2:], clouds[2:]):
        assert r.error is None and np.array_equal(r.labels.numpy(), oracle.dbscan2d(c, 0.4, 6))


def _wide_key_cloud(seed, n=30_000):
    """Points over [-3e5, 3e5]^2 with eps 0.001: about 30-bit cell coordinates
    each way, so 60-bit keys and more than four radix passes; tight groups of
    ten points make clusters."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-3e5, 3e5, (n // 20, 2))
    groups = np.repeat(centres, 10, axis=0) + rng.uniform(-3e-4, 3e-4, (len(centres) * 10, 2))
    far = rng.uniform(-3e5, 3e5, (n - len(groups), 2))
    X = np.concatenate([groups, far]).astype(np.float32)
    return X[rng.permutation(len(X))]


def test_wide_key_batches_on_two_streams(dt):
    """Two batches of wide-key clouds (their sort runs the ticketed tail passes)
    enqueued on two streams at once, three times: every cloud equals the
    oracle."""
    batches = [[_wide_key_cloud(100 * b + i) for i in range(4)] for b in range(2)]
    dbg = dt._C.dbscan2d_hip_debug(torch.from_numpy(batches[0][0]).to(DEV), 0.001, 3)
    assert int(dbg[1][5].item()) > 4  # radix passes beyond the fourth
    refs = [[oracle.dbscan2d(c, 0.001, 3) for c in b] for b in batches]
    inputs = [[torch.from_numpy(c).to(DEV) for c in b] for b in batches]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for rep in range(3):
        print(f"[gpu errors] two-stream wide-key batches, round {rep}", flush=True)
        outs = []
        for s, b in zip(streams, inputs):
            with torch.cuda.stream(s):
                outs.append(dt.dbscan2d_batched(b, 0.001, 3))
        torch.cuda.synchronize()
        for o, r in zip(outs, refs):
            for oi, ri in zip(o, r):
                assert np.array_equal(oi.cpu().numpy(), ri)
    dt.check_device_errors()
