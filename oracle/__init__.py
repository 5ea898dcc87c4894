"""Test-only oracle for the 2-D grid DBSCAN hot path.

TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package, and only as the checker.  The product
(dbscan-torch_amd/) never imports it.

Two CPU implementations live here:

* ``dbscan2d`` / ``dbscan2d_debug`` -- the C restatement in dbscan_oracle.c of
  the reference CPU path (csrc/cpu/dbscan_grid.hpp:136-458,
  csrc/cpu/dbscan_cpu.cpp:34-59), loaded from liboracle.so through ctypes.
* ``ref_dbscan2d`` -- the reference's own CPU path compiled from its sources by
  ``make -C oracle ref`` into oracle/_ref/_C.so (a pybind module named ``_C``).
  Optional: present only where the reference tree was available at build time.
"""

from __future__ import annotations

import ctypes
import importlib.machinery
import importlib.util
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "_C.so")

_lib = None
_ref = None


def build(ref: bool = True) -> None:
    """Compile liboracle.so, and _ref/_C.so when the reference tree exists."""
    subprocess.run(["make", "-C", _HERE, "-s", "liboracle.so"], check=True)
    ref_dir = os.environ.get("DBSCAN_REF_DIR", "/root/reference")
    if ref and os.path.isdir(os.path.join(ref_dir, "csrc")):
        subprocess.run(["make", "-C", _HERE, "-s", "ref", f"REF_DIR={ref_dir}"], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        lib.oracle_dbscan2d.argtypes = [P, ctypes.c_int64, ctypes.c_double, ctypes.c_int64, P]
        lib.oracle_dbscan2d.restype = ctypes.c_int
        lib.oracle_dbscan2d_debug.argtypes = [P, ctypes.c_int64, ctypes.c_double, ctypes.c_int64,
                                              P, P, P, P, P, P, P, P, P, P, P]
        lib.oracle_dbscan2d_debug.restype = ctypes.c_int
        lib.oracle_set_threads.argtypes = [ctypes.c_int]
        lib.oracle_set_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def _as_points(X) -> np.ndarray:
    X = np.ascontiguousarray(np.asarray(X), dtype=np.float32)
    if X.ndim != 2 or X.shape[1] != 2:
        raise ValueError("X must have shape (N, 2)")
    return X


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(threads: int) -> int:
    """Set the OpenMP thread count used by the C oracle; returns the count in effect."""
    return int(_load().oracle_set_threads(int(threads)))


def dbscan2d(X, eps: float, min_samples: int) -> np.ndarray:
    """Labels (int32, input order, -1 = noise) from the C restatement."""
    X = _as_points(X)
    n = X.shape[0]
    labels = np.empty(n, dtype=np.int32)
    rc = _load().oracle_dbscan2d(_ptr(X), n, float(eps), int(min_samples), _ptr(labels))
    if rc != 0:
        raise RuntimeError(f"oracle_dbscan2d failed with code {rc}")
    return labels


def dbscan2d_debug(X, eps: float, min_samples: int) -> dict:
    """Labels plus the stage intermediates of the C restatement, in the
    reference's cell order: perm, is_core (sorted order), cell_start, cell_kx,
    cell_ky, cell_status, cell_root, cell_cluster, num_cells, n_clusters."""
    X = _as_points(X)
    n = X.shape[0]
    m = max(n, 1)
    out = {
        "labels": np.empty(n, dtype=np.int32),
        "perm": np.empty(n, dtype=np.int32),
        "is_core": np.empty(n, dtype=np.uint8),
        "cell_start": np.empty(m + 1, dtype=np.int32),
        "cell_kx": np.empty(m, dtype=np.int32),
        "cell_ky": np.empty(m, dtype=np.int32),
        "cell_status": np.empty(m, dtype=np.uint8),
        "cell_root": np.empty(m, dtype=np.int32),
        "cell_cluster": np.empty(m, dtype=np.int32),
    }
    nc = ctypes.c_int32(0)
    ncl = ctypes.c_int32(0)
    rc = _load().oracle_dbscan2d_debug(
        _ptr(X), n, float(eps), int(min_samples), _ptr(out["labels"]), _ptr(out["perm"]),
        _ptr(out["is_core"]), _ptr(out["cell_start"]), _ptr(out["cell_kx"]), _ptr(out["cell_ky"]),
        _ptr(out["cell_status"]), _ptr(out["cell_root"]), _ptr(out["cell_cluster"]),
        ctypes.byref(nc), ctypes.byref(ncl))
    if rc != 0:
        raise RuntimeError(f"oracle_dbscan2d_debug failed with code {rc}")
    c = nc.value
    out["num_cells"] = c
    out["n_clusters"] = ncl.value
    out["cell_start"] = out["cell_start"][: c + 1]
    for k in ("cell_kx", "cell_ky", "cell_status", "cell_root", "cell_cluster"):
        out[k] = out[k][:c]
    return out


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def _load_ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference build missing: {_REF_PATH} (run make -C oracle ref)")
        import torch  # noqa: F401  -- the module links libtorch; load it first

        loader = importlib.machinery.ExtensionFileLoader("_C", _REF_PATH)
        spec = importlib.util.spec_from_file_location("_C", _REF_PATH, loader=loader)
        mod = importlib.util.module_from_spec(spec)
        loader.exec_module(mod)
        _ref = mod
    return _ref


def ref_dbscan2d(X, eps: float, min_samples: int) -> np.ndarray:
    """Labels from the reference's own compiled CPU path (oracle/_ref/_C.so)."""
    import torch

    Xt = torch.from_numpy(_as_points(X))
    return _load_ref().dbscan2d(Xt, float(eps), int(min_samples)).numpy()


def ref_dbscan2d_profile(X, eps: float, min_samples: int):
    """(labels, {stage: ms}) from the reference's compiled CPU path."""
    import torch

    Xt = torch.from_numpy(_as_points(X))
    labels, stages = _load_ref().dbscan2d_profile(Xt, float(eps), int(min_samples))
    return labels.numpy(), dict(stages)
