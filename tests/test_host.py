"""Host-side checks that need no GPU: the binding's validation (same order and
messages as the reference, csrc/extension.cpp:26-31), the C-ABI library's
exports and its argument checks, which return before any device call."""

import ctypes
import os
import re

import numpy as np
import pytest
import torch

import dbscan_torch
from dbscan_torch import DBSCAN, dbscan2d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dbscan_hip.h")
LIB = os.path.join(os.path.dirname(dbscan_torch.__file__), "libdbscan_hip.so")


def test_input_validation_messages():
    # tests/test_basic.py:67-75 of the reference, same patterns
    with pytest.raises(RuntimeError, match=r"shape \(N, 2\)"):
        dbscan2d(torch.zeros((10, 3)), eps=1.0, min_samples=5)
    with pytest.raises(RuntimeError, match="float32"):
        dbscan2d(torch.zeros((10, 2), dtype=torch.float64), eps=1.0, min_samples=5)
    with pytest.raises(RuntimeError, match="eps"):
        dbscan2d(torch.zeros((10, 2)), eps=0.0, min_samples=5)
    with pytest.raises(RuntimeError, match="min_samples"):
        dbscan2d(torch.zeros((10, 2)), eps=1.0, min_samples=0)


def test_validation_order():
    # shape is checked before dtype, dtype before eps, eps before min_samples
    with pytest.raises(RuntimeError, match=r"shape \(N, 2\)"):
        dbscan2d(torch.zeros((4, 3), dtype=torch.float64), eps=-1.0, min_samples=0)
    with pytest.raises(RuntimeError, match="float32"):
        dbscan2d(torch.zeros((4, 2), dtype=torch.float64), eps=-1.0, min_samples=0)
    with pytest.raises(RuntimeError, match="eps"):
        dbscan2d(torch.zeros((4, 2)), eps=float("nan"), min_samples=0)
    with pytest.raises(RuntimeError, match="eps"):
        dbscan2d(torch.zeros((4, 2)), eps=-1.0, min_samples=0)


def test_empty_input_needs_no_device():
    X = torch.empty((0, 2), dtype=torch.float32)
    labels = dbscan2d(X, eps=1.0, min_samples=5)
    assert labels.shape == (0,) and labels.dtype == torch.int32 and labels.device == X.device
    labels, stages = dbscan_torch._C.dbscan2d_profile(X, 1.0, 5)
    assert labels.shape == (0,)
    assert list(stages) == sorted(["cell_keys", "sort_gather", "boundary_scan", "build_ngh",
                                   "core_mark", "union_find", "cluster_id", "core_labels",
                                   "border_assign", "unpermute"])


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-device error path")
def test_host_tensor_without_device_raises():
    X = torch.zeros((10, 2))
    with pytest.raises(RuntimeError, match="no HIP device"):
        dbscan2d(X, eps=1.0, min_samples=5)
    with pytest.raises(RuntimeError, match="no HIP device"):
        DBSCAN(eps=1.0, min_samples=5).fit(np.zeros((10, 2), np.float32))


def _declared_functions():
    text = open(HEADER).read()
    return re.findall(r"^(?:int|void|const char\*)\s+(dbscan2d_\w+)\s*\(", text, re.M)


def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert {"dbscan2d_hip", "dbscan2d_hip_workspace_bytes", "dbscan2d_hip_profile"} <= set(names)
    lib = ctypes.CDLL(LIB)
    for name in names:
        assert hasattr(lib, name), name


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    lib.dbscan2d_hip.argtypes = [P, ctypes.c_int64, ctypes.c_double, ctypes.c_int64, P, P,
                                 ctypes.c_size_t, P]
    lib.dbscan2d_hip.restype = ctypes.c_int
    lib.dbscan2d_hip_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]
    lib.dbscan2d_hip_status_string.restype = ctypes.c_char_p
    lib.dbscan2d_hip_stage_name.restype = ctypes.c_char_p
    return lib


def test_c_abi_argument_checks(lib):
    # codes from include/dbscan_hip.h; each returns before any device work
    assert lib.dbscan2d_hip(None, -1, 1.0, 5, None, None, 0, None) == 1
    assert lib.dbscan2d_hip(None, 2**31, 1.0, 5, None, None, 0, None) == 1
    assert lib.dbscan2d_hip(None, 10, 0.0, 5, None, None, 0, None) == 2
    assert lib.dbscan2d_hip(None, 10, float("nan"), 5, None, None, 0, None) == 2
    assert lib.dbscan2d_hip(None, 10, 1.0, 0, None, None, 0, None) == 3
    assert lib.dbscan2d_hip(None, 0, 1.0, 5, None, None, 0, None) == 0
    assert lib.dbscan2d_hip(None, 10, 1.0, 5, None, None, 0, None) == 4
    # misaligned workspace, then a too-small one (fake addresses, never touched)
    assert lib.dbscan2d_hip(0x1000, 10, 1.0, 5, 0x2000, 0x3001, 1 << 30, None) == 6
    assert lib.dbscan2d_hip(0x1000, 10, 1.0, 5, 0x2000, 0x3000, 16, None) == 5
    assert lib.dbscan2d_hip_status_string(5) == b"workspace too small"


def test_workspace_and_layout(lib):
    sizes = []
    for n in (1, 1000, 4096, 4097, 10_000_000):
        b = ctypes.c_size_t(0)
        assert lib.dbscan2d_hip_workspace_bytes(n, ctypes.byref(b)) == 0
        sizes.append(b.value)
    assert sizes == sorted(sizes)
    # about 110 bytes per point at scale
    assert 100 * 10_000_000 < sizes[-1] < 130 * 10_000_000
    names = [lib.dbscan2d_hip_stage_name(i).decode() for i in range(10)]
    assert names[0] == "cell_keys" and names[-1] == "unpermute"
    assert lib.dbscan2d_hip_stage_name(10) is None


def test_class_wrapper_contract_without_compute():
    clf = DBSCAN()
    assert (clf.eps, clf.min_samples, clf.labels_) == (0.5, 5, None)
    out = DBSCAN(eps=1.0, min_samples=3).fit(np.zeros((0, 2), np.float32))
    assert isinstance(out.labels_, np.ndarray) and out.labels_.shape == (0,)


def test_batched_validation_messages():
    from dbscan_torch import dbscan2d_batched

    a, b = torch.zeros((4, 2)), torch.zeros((3, 2))
    with pytest.raises(RuntimeError, match=r"eps must be > 0 \(cloud 1\)"):
        dbscan2d_batched([a, b], [0.5, 0.0], 5)
    with pytest.raises(RuntimeError, match=r"min_samples must be >= 1 \(cloud 0\)"):
        dbscan2d_batched([a, b], 0.5, [0, 5])
    with pytest.raises(RuntimeError, match=r"shape \(N, 2\)"):
        dbscan2d_batched([a, torch.zeros((3, 3))], 0.5, 5)
    with pytest.raises(RuntimeError, match="float32"):
        dbscan2d_batched([a, b.double()], 0.5, 5)
    with pytest.raises(ValueError, match="2 values for 3 clouds"):
        dbscan2d_batched([a, b, a], [0.5, 0.5], 5)
    packed = torch.zeros((7, 2))
    with pytest.raises(RuntimeError, match="offsets must run from 0 to N"):
        dbscan2d_batched(packed, 0.5, 5, offsets=[0, 4, 6])
    with pytest.raises(RuntimeError, match="non-decreasing"):
        dbscan2d_batched(packed, 0.5, 5, offsets=[0, 5, 4, 7])
    assert dbscan2d_batched([], 0.5, 5) == []
    out = dbscan2d_batched(torch.zeros((0, 2)), 0.5, 5, offsets=[0, 0, 0])
    assert out.shape == (0,) and out.dtype == torch.int32


def test_batched_c_abi_argument_checks():
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    lib.dbscan2d_hip_batched.argtypes = [P, ctypes.c_int64, P, P, P, ctypes.c_int64, P, P,
                                         ctypes.c_size_t, P, ctypes.POINTER(ctypes.c_int), P]
    lib.dbscan2d_hip_batched_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int64,
                                                         ctypes.POINTER(ctypes.c_size_t)]
    offs = np.array([0, 6, 6, 10], dtype=np.int64)
    eps = np.array([0.5, 1.0, -1.0], dtype=np.float64)
    ms = np.array([5, 0, 5], dtype=np.int64)
    bad = ctypes.c_int(0)

    def call(o, e, m, total=10, B=3, ws=0x3000, ws_bytes=1 << 30):
        return lib.dbscan2d_hip_batched(0x1000, total, o.ctypes.data, e.ctypes.data,
                                        m.ctypes.data, B, 0x2000, ws, ws_bytes, None,
                                        ctypes.byref(bad), None)

    assert call(offs, eps, ms) == 3 and bad.value == 1          # min_samples of cloud 1
    ms[1] = 5
    assert call(offs, eps, ms) == 2 and bad.value == 2          # eps of cloud 2
    eps[2] = 2.0
    assert call(offs, eps, ms, total=11) == 1                   # offsets do not end at N
    assert call(np.array([0, 7, 6, 10], dtype=np.int64), eps, ms) == 1 and bad.value == 1
    assert call(offs, eps, ms, ws=0x3001) == 6                  # misaligned workspace
    assert call(offs, eps, ms, ws_bytes=16) == 5                # workspace too small
    sizes = []
    for total, B in ((10, 1), (10, 3), (1_000_000, 1), (1_000_000, 2048)):
        b = ctypes.c_size_t(0)
        assert lib.dbscan2d_hip_batched_workspace_bytes(total, B, ctypes.byref(b)) == 0
        sizes.append(b.value)
    assert sizes[0] <= sizes[1] < sizes[2] < sizes[3]


class _Cfg(ctypes.Structure):
    _fields_ = [("sort_items", ctypes.c_int32), ("radix_bits", ctypes.c_int32),
                ("cell_blocks", ctypes.c_int32), ("pair_blocks", ctypes.c_int32),
                ("small_max", ctypes.c_int32)]


def test_launch_configuration_checks():
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    lib.dbscan2d_hip_configured.argtypes = [P, ctypes.c_int64, ctypes.c_double, ctypes.c_int64,
                                            P, P, ctypes.c_size_t, P, ctypes.POINTER(_Cfg)]
    lib.dbscan2d_hip_default_config.argtypes = [ctypes.c_int64, ctypes.POINTER(_Cfg)]
    d = _Cfg()
    assert lib.dbscan2d_hip_default_config(10_000_000, ctypes.byref(d)) == 0
    assert d.sort_items in (8, 12, 16) and 4 <= d.radix_bits <= 9
    assert d.small_max == 4096  # DBSCAN_HIP_SMALL_MAX
    for bad in ((10, 8, 8192, 8192), (12, 10, 8192, 8192), (12, 3, 8192, 8192),
                (12, 8, 10, 8192), (12, 8, 8192, 1 << 20), (12, 8, 8192, 8192, -1),
                (12, 8, 8192, 8192, 4097)):
        c = _Cfg(*bad)
        # fake, aligned addresses: the configuration is refused before any device work
        assert lib.dbscan2d_hip_configured(0x1000, 10, 1.0, 5, 0x2000, 0x3000, 1 << 30, None,
                                           ctypes.byref(c)) == 8, bad


def test_tuning_table_round_trip():
    from dbscan_torch import _C, autotune

    saved = _C.get_tuning_table()
    try:
        autotune.install_rows([{"max_n": 50_000, "config": [8, 7, 2048, 4096]},
                               {"max_n": 2**31 - 1, "config": [16, 8, 8192, 16384]}])
        # rows without small_max take the library's default (4096)
        assert _C.get_tuning_table() == [[50_000, 8, 7, 2048, 4096, 4096],
                                         [2**31 - 1, 16, 8, 8192, 16384, 4096]]
        autotune.install_rows([{"max_n": 50_000, "config": [8, 7, 2048, 4096, 1000]}])
        assert _C.get_tuning_table() == [[50_000, 8, 7, 2048, 4096, 1000]]
        with pytest.raises(RuntimeError, match="tuning rows"):
            _C.set_tuning_table([[1, 2, 3]])
    finally:
        _C.set_tuning_table(saved)
    assert _C.default_config(1000) == [8, 8, 2048, 16384, 4096]


def test_synthetic_tuning_cloud_is_seeded():
    from dbscan_torch.autotune import synthetic_cloud

    a, b = synthetic_cloud(5000, 3), synthetic_cloud(5000, 3)
    assert a.shape == (5000, 2) and a.dtype == torch.float32 and torch.equal(a, b)
    assert not torch.equal(a, synthetic_cloud(5000, 4))


def test_device_error_bits_map_to_runtime_errors(lib):
    """Every device-side error bit the header declares has a name, and a
    non-zero word raises RuntimeError naming each of its bits (host logic
    only: no device call)."""
    text = open(HEADER).read()
    bits = {name: int(v) for name, v in re.findall(r"#define DBSCAN_HIP_DEV_(\w+) (\d+)", text)}
    assert set(bits) == {"SORT_WAIT", "SCAN_WAIT", "UF_LOOP", "RANGE"}
    lib.dbscan2d_hip_device_error_name.restype = ctypes.c_char_p
    lib.dbscan2d_hip_device_error_name.argtypes = [ctypes.c_int32]
    names = {}
    for key, bit in bits.items():
        name = lib.dbscan2d_hip_device_error_name(bit)
        assert name, key
        names[bit] = name.decode()
        assert names[bit] in dbscan_torch._C.device_error_message(bit)
        with pytest.raises(RuntimeError, match=re.escape(names[bit])):
            dbscan_torch._C.raise_for_device_error(bit)
    assert lib.dbscan2d_hip_device_error_name(16) is None
    word = sum(bits.values())
    msg = dbscan_torch._C.device_error_message(word)
    assert all(n in msg for n in names.values()) and "0xf" in msg and "INT32_MIN" in msg
    dbscan_torch._C.raise_for_device_error(0)  # a clean word raises nothing
    lib.dbscan2d_hip_status_string.argtypes = [ctypes.c_int]
    assert b"device-side check failed" in lib.dbscan2d_hip_status_string(9)


def test_error_readback_argument_checks(lib):
    P = ctypes.c_void_p
    lib.dbscan2d_hip_error_async.argtypes = [P, P, P]
    lib.dbscan2d_hip_check.argtypes = [P, P, P]
    word = ctypes.c_int32(0)
    assert lib.dbscan2d_hip_error_async(None, ctypes.addressof(word), None) == 4
    assert lib.dbscan2d_hip_error_async(0x1000, None, None) == 4
    assert lib.dbscan2d_hip_error_async(0x1001, ctypes.addressof(word), None) == 6
    assert lib.dbscan2d_hip_check(None, None, None) == 4
    assert lib.dbscan2d_hip_check(0x1010, None, ctypes.addressof(word)) == 6


def test_check_device_errors_without_calls():
    dbscan_torch.check_device_errors()  # nothing in flight: returns
