"""dbscan_torch on MI355X: 2-D grid-hashed DBSCAN for PyTorch-ROCm.

Drop-in for the reference package's public surface
(dbscan_torch/__init__.py:45-118): ``dbscan2d``, ``DBSCAN``, ``ArrayLike2D``
and ``__version__``, with the same argument meaning, input/output types and
error behaviour.  The clustering runs in hand-written gfx950 HIP kernels
(libdbscan_hip.so) reached through the pybind module ``_C``; labels are
bit-identical to the reference's CPU path.
"""

from importlib.metadata import PackageNotFoundError
from importlib.metadata import version as _pkg_version
from typing import Any, Optional, Sequence, Union

import numpy as np
import torch

# Same guard as the reference (dbscan_torch/__init__.py:11-38): a binding built
# against another torch minor version fails to load with an obscure symbol
# error, so name the mismatch first.
try:
    from ._build_info import BUILD_TORCH_VERSION as _build_torch
except ImportError:
    _build_torch = None

if _build_torch is not None:
    try:
        _mismatch = tuple(int(x) for x in _build_torch.split(".")[:2]) != tuple(
            int(x) for x in torch.__version__.split(".")[:2])
    except (ValueError, IndexError, AttributeError):
        _mismatch = False
    if _mismatch:
        raise ImportError(
            f"dbscan_torch was built against torch {_build_torch} but the installed runtime "
            f"torch is {torch.__version__}. Rebuild with: python dbscan-torch_amd/build.py")

try:
    from . import _C
except ImportError as exc:  # the native library is the product: fail loudly
    raise ImportError(
        "dbscan_torch: the native extension (_C, libdbscan_hip.so) is not built or does not "
        "load. Build it with: python dbscan-torch_amd/build.py") from exc

# Launch configurations saved by the autotuner (python -m dbscan_torch.autotune)
# beside this file; without one the library's built-in table applies.
def _load_tuning() -> None:
    import json
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.json")
    if os.path.exists(path):
        with open(path) as f:
            rows = json.load(f)["rows"]
        _C.set_tuning_table([[int(r["max_n"])] + [int(v) for v in r["config"]] for r in rows])


_load_tuning()

ArrayLike2D = Union[np.ndarray, torch.Tensor]

__all__ = ["dbscan2d", "dbscan2d_batched", "BatchedGraph", "DBSCAN", "ArrayLike2D",
           "check_device_errors"]

try:
    __version__ = _pkg_version("dbscan-torch")
except PackageNotFoundError:
    __version__ = "0.2.0+mi355x"


def dbscan2d(X: torch.Tensor, eps: float, min_samples: int) -> torch.Tensor:
    """Cluster the rows of a float32 ``(N, 2)`` tensor with grid-hashed DBSCAN.

    ``eps`` is the neighbourhood radius (must be > 0); a point is a core point
    when at least ``min_samples`` points (itself included) lie within
    ``eps``.  The result is an int32 ``(N,)`` tensor on X's device: ``-1`` for
    noise, else a cluster id in ``[0, n_clusters)`` numbered as the
    reference CPU path numbers them.  A ROCm tensor is clustered where it
    lies and the call returns without waiting for the GPU; a host tensor is
    copied to the current HIP device and its labels come back to the host.

    A failed device-side check (never expected) leaves every label INT32_MIN
    and raises RuntimeError: at once for a host tensor, else from the next
    call on the device or from :func:`check_device_errors`.
    """
    return _C.dbscan2d(X, float(eps), int(min_samples))


def check_device_errors(device: Optional[int] = None) -> None:
    """Wait for every dbscan2d / dbscan2d_batched call that returned device
    labels without waiting, and raise RuntimeError if one of them failed a
    device-side check (a bounded wait that gave up, an index out of range)."""
    _C.check_device_errors(device)


def _per_cloud(value, count: int, name: str, cast):
    if isinstance(value, (list, tuple, np.ndarray, torch.Tensor)):
        vals = [cast(v) for v in (value.tolist() if hasattr(value, "tolist") else value)]
        if len(vals) != count:
            raise ValueError(f"dbscan2d_batched: {name} has {len(vals)} values for {count} clouds")
        return vals
    return [cast(value)] * count


def dbscan2d_batched(clouds: Union[Sequence[torch.Tensor], torch.Tensor],
                     eps: Union[float, Sequence[float]],
                     min_samples: Union[int, Sequence[int]],
                     offsets: Optional[Sequence[int]] = None):
    """Cluster many independent 2-D point clouds in one batched launch.

    No counterpart in the reference (which clusters one cloud per call); every
    cloud's labels are exactly those ``dbscan2d(cloud, eps_i, min_samples_i)``
    returns, numbered from 0 within the cloud.

    Args:
        clouds: a sequence of float32 ``(N_i, 2)`` tensors (all on one device),
            or one packed ``(N, 2)`` tensor together with ``offsets``.
        eps: one radius for every cloud, or one per cloud (each > 0).
        min_samples: one value for every cloud, or one per cloud (each >= 1).
        offsets: with a packed tensor, ``B + 1`` row offsets from 0 to N; cloud
            i is rows ``[offsets[i], offsets[i+1])``.

    Returns:
        With a sequence: a list of int32 label tensors, one per cloud, on the
        input device.  With a packed tensor: the packed int32 labels ``(N,)``.
    """
    if isinstance(clouds, torch.Tensor):
        if offsets is None:
            raise ValueError("dbscan2d_batched: a packed tensor needs offsets")
        offs = [int(v) for v in (offsets.tolist() if hasattr(offsets, "tolist") else offsets)]
        B = len(offs) - 1
        return _C.dbscan2d_batched(clouds, offs, _per_cloud(eps, B, "eps", float),
                                   _per_cloud(min_samples, B, "min_samples", int))
    parts = list(clouds)
    B = len(parts)
    if B == 0:
        return []
    for p in parts:
        if not (isinstance(p, torch.Tensor) and p.dim() == 2 and p.size(1) == 2):
            raise RuntimeError("X must have shape (N, 2)")
        if p.dtype != torch.float32:
            raise RuntimeError("X must be float32")
    sizes = [int(p.size(0)) for p in parts]
    offs = [0]
    for n in sizes:
        offs.append(offs[-1] + n)
    packed = torch.cat(parts, 0) if B > 1 else parts[0]
    labels = _C.dbscan2d_batched(packed, offs, _per_cloud(eps, B, "eps", float),
                                 _per_cloud(min_samples, B, "min_samples", int))
    return list(torch.split(labels, sizes))


class BatchedGraph:
    """A batched dbscan2d captured in a HIP graph once and replayed.

    No counterpart in the reference.  ``X`` is one packed float32 ``(N, 2)``
    tensor on a ROCm device and ``offsets``, ``eps`` and ``min_samples`` are as
    for :func:`dbscan2d_batched`.  The batch's tables are staged into a
    workspace of the object's own, one eager run checks the set-up, and the
    kernels (no copy, allocation or host synchronisation among them) are
    captured in a ``torch.cuda.CUDAGraph``.  The graph reads ``X`` and writes
    :attr:`labels` in place, so refill ``X`` with other clouds of the same
    sizes and call :meth:`replay` for their labels; :meth:`restage` changes
    eps and min_samples for the same sizes.  Runs of one object must not
    overlap (they share the workspace).
    """

    def __init__(self, X: torch.Tensor, offsets: Sequence[int],
                 eps: Union[float, Sequence[float]], min_samples: Union[int, Sequence[int]]):
        if not (isinstance(X, torch.Tensor) and X.is_cuda):
            raise RuntimeError("BatchedGraph: X must be a tensor on a ROCm device")
        self.X = X
        self.offsets = [int(v) for v in (offsets.tolist() if hasattr(offsets, "tolist")
                                         else offsets)]
        B = len(self.offsets) - 1
        dev = X.device
        self.workspace = torch.empty(_C.dbscan2d_batched_workspace_bytes(int(X.size(0)), B),
                                     dtype=torch.uint8, device=dev)
        self.labels = torch.empty(int(X.size(0)), dtype=torch.int32, device=dev)
        self.error_word = torch.zeros(1, dtype=torch.int32, device=dev)
        self.plan = None
        self.restage(eps, min_samples)
        # one eager run: kernels loaded, arguments checked on the device
        _C.dbscan2d_batched_enqueue(self.X, self.labels, self.workspace, self.plan,
                                    self.error_word)
        self.check()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            _C.dbscan2d_batched_enqueue(self.X, self.labels, self.workspace, self.plan,
                                        self.error_word)

    def restage(self, eps: Union[float, Sequence[float]],
                min_samples: Union[int, Sequence[int]]) -> None:
        """Stages eps and min_samples (the cloud sizes stay those of the
        capture); the next replay uses them."""
        B = len(self.offsets) - 1
        plan = _C.dbscan2d_batched_stage(self.X, self.offsets, _per_cloud(eps, B, "eps", float),
                                         _per_cloud(min_samples, B, "min_samples", int),
                                         self.workspace)
        if self.plan is not None and plan != self.plan:
            raise RuntimeError("BatchedGraph.restage: the staged launch plan changed; "
                               "capture a new BatchedGraph")
        self.plan = plan
        torch.cuda.current_stream(self.X.device).synchronize()

    def replay(self) -> torch.Tensor:
        """Runs the captured batch on the current contents of X; returns
        :attr:`labels` (packed int32, written in stream order)."""
        self.graph.replay()
        return self.labels

    def check(self) -> None:
        """Waits for the last run and raises RuntimeError if it failed a
        device-side check."""
        bits = int(self.error_word.item())
        if bits:
            _C.raise_for_device_error(bits)


class DBSCAN:
    """Estimator-style front end to :func:`dbscan2d` with the reference's
    constructor defaults and methods (dbscan_torch/__init__.py:76-118).

    ``fit`` stores the labels in ``labels_`` in the container type it was
    given (a numpy array for numpy input, a tensor on the input's device for
    a tensor); input of another dtype is converted to float32 first.
    """

    def __init__(self, eps: float = 0.5, min_samples: int = 5):
        self.eps = float(eps)
        self.min_samples = int(min_samples)
        self.labels_: Optional[ArrayLike2D] = None

    def fit(self, X: ArrayLike2D, y: Any = None) -> "DBSCAN":
        """Cluster ``X`` and keep its labels in ``labels_`` (``y`` is unused)."""
        is_torch = isinstance(X, torch.Tensor)
        if is_torch:
            Xt = X if X.dtype == torch.float32 else X.to(torch.float32)
        else:
            Xt = torch.as_tensor(X, dtype=torch.float32)
        labels = dbscan2d(Xt, self.eps, self.min_samples)
        self.labels_ = labels if is_torch else labels.cpu().numpy()
        return self

    def fit_predict(self, X: ArrayLike2D, y: Any = None) -> ArrayLike2D:
        """``fit(X)`` followed by returning ``labels_``."""
        return self.fit(X, y).labels_  # type: ignore[return-value]
