"""Writes the golden label fixtures of tests/golden/ from the reference's own
CPU path, compiled from its sources into oracle/_ref/_C.so by
``make -C oracle ref`` (see oracle/Makefile).

Run in the build container (where the reference tree exists):

    python tests/golden/make_golden.py

Outputs (all data, no code):
  blobs10k_labels.npy   int32 labels of the 10,000-point workload, input order
  cases.npz         small constructed inputs and their reference labels
                    (x_<name>, eps_<name>, ms_<name>, labels_<name>)
  golden.json       sha256 of the int32 label bytes, cluster count, noise
                    count and cell count for blobs10k, lattice1m, blobs10m and skewed10m
"""

from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from tests import workloads as W  # noqa: E402
from tests.golden.cases import constructed_cases  # noqa: E402


def summary(X: np.ndarray, eps: float, ms: int, labels: np.ndarray) -> dict:
    kx, ky = W.cell_keys_np(X, eps)
    cells = np.unique((kx << 32) | (ky & 0xFFFFFFFF)).size
    return {
        "n": int(X.shape[0]), "eps": float(eps), "min_samples": int(ms),
        "sha256": hashlib.sha256(np.ascontiguousarray(labels, dtype=np.int32).tobytes()).hexdigest(),
        "n_clusters": int(labels.max() + 1) if (labels >= 0).any() else 0,
        "n_noise": int((labels == -1).sum()), "n_cells": int(cells),
    }


def main() -> None:
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/_C.so is not built: run make -C oracle ref")
    golden = {}
    for name, fn in W.WORKLOADS.items():
        X, eps, ms = fn()
        ref = oracle.ref_dbscan2d(X, eps, ms)
        golden[name] = summary(X, eps, ms, ref)
        if name == "blobs10k":
            np.save(os.path.join(HERE, "blobs10k_labels.npy"), ref.astype(np.int32))
        print(name, golden[name], flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(golden, f, indent=2, sort_keys=True)

    arrays = {}
    for name, (X, eps, ms) in constructed_cases().items():
        arrays[f"x_{name}"] = X
        arrays[f"eps_{name}"] = np.float64(eps)
        arrays[f"ms_{name}"] = np.int64(ms)
        arrays[f"labels_{name}"] = oracle.ref_dbscan2d(X, eps, ms).astype(np.int32)
    np.savez(os.path.join(HERE, "cases.npz"), **arrays)
    print("cases:", sorted(constructed_cases()))


if __name__ == "__main__":
    main()
