"""Writes the golden label fixtures of tests/golden/ from the reference's own
CPU path, compiled from its sources into oracle/_ref/_C.so by
``make -C oracle ref`` (see oracle/Makefile).

Run in the build container (where the reference tree exists):

    python tests/golden/make_golden.py            # everything
    python tests/golden/make_golden.py streams    # streams.json only

Outputs (all data, no code):
  blobs10k_labels.npy   int32 labels of the 10,000-point workload, input order
  cases.npz         small constructed inputs and their reference labels
                    (x_<name>, eps_<name>, ms_<name>, labels_<name>)
  golden.json       sha256 of the int32 label bytes, cluster count, noise
                    count and cell count for blobs10k, lattice1m, blobs10m and skewed10m
  streams.json      per-cloud label digests of the batched and served sets:
                    "clouds2048" -- one digest per cloud of the 2,048-cloud
                    batch (tests.workloads.clouds2048_*), in cloud order;
                    "serve" -- for each rate of SERVE_PINNED_RATES, the first
                    SERVE_PINNED_COUNT requests of that stream (N, eps,
                    min_samples and a digest each; points from
                    serve_request_cloud).  A digest is the first 16 hex
                    digits of the sha256 of the int32 label bytes.
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


def digest(labels: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(labels, dtype=np.int32).tobytes()).hexdigest()[:16]


def write_streams() -> None:
    out = {"digest": "sha256(int32 label bytes)[:16]"}
    ns, eps, ms = W.clouds2048_params()
    hashes = []
    for i in range(len(ns)):
        X = W.clouds2048_cloud(i, int(ns[i]))
        hashes.append(digest(oracle.ref_dbscan2d(X, float(eps[i]), int(ms[i]))))
        if i % 256 == 255:
            print(f"clouds2048: {i + 1} clouds", flush=True)
    out["clouds2048"] = {"points": int(ns.sum()), "digests": hashes}
    out["serve"] = {}
    for rate in W.SERVE_PINNED_RATES:
        _, rns, reps, rms = W.serve_stream(rate, W.SERVE_PINNED_COUNT)
        reqs = []
        for i in range(W.SERVE_PINNED_COUNT):
            X = W.serve_request_cloud(rate, i, int(rns[i]))
            lab = oracle.ref_dbscan2d(X, float(reps[i]), int(rms[i]))
            reqs.append([int(rns[i]), float(reps[i]), int(rms[i]), digest(lab)])
        out["serve"][str(rate)] = {"points": int(rns.sum()), "requests": reqs}
        print(f"serve {rate}: {W.SERVE_PINNED_COUNT} requests, {int(rns.sum())} points", flush=True)
    with open(os.path.join(HERE, "streams.json"), "w") as f:
        json.dump(out, f, indent=1)


def main() -> None:
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/_C.so is not built: run make -C oracle ref")
    if sys.argv[1:] == ["streams"]:
        write_streams()
        return
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
    write_streams()


if __name__ == "__main__":
    main()
